// Small vector math and codecs used by the gfx950 kernels.
//
// Everything here is plain IEEE FP32 arithmetic: the library is compiled
// with -ffp-contract=off, so a*b+c is never fused and the kernels round
// exactly like a scalar CPU evaluation of the same expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace rg
{

#define RG_HD __host__ __device__ __forceinline__

struct v3
{
    float x, y, z;
};

RG_HD v3 mk3( float x, float y, float z )
{
    return v3{ x, y, z };
}
RG_HD v3 operator+( v3 a, v3 b )
{
    return v3{ a.x + b.x, a.y + b.y, a.z + b.z };
}
RG_HD v3 operator-( v3 a, v3 b )
{
    return v3{ a.x - b.x, a.y - b.y, a.z - b.z };
}
RG_HD v3 operator-( v3 a )
{
    return v3{ -a.x, -a.y, -a.z };
}
RG_HD v3 operator*( v3 a, float s )
{
    return v3{ a.x * s, a.y * s, a.z * s };
}
RG_HD v3 operator*( float s, v3 a )
{
    return v3{ s * a.x, s * a.y, s * a.z };
}
RG_HD v3 operator*( v3 a, v3 b )
{
    return v3{ a.x * b.x, a.y * b.y, a.z * b.z };
}
RG_HD v3 operator/( v3 a, float s )
{
    return v3{ a.x / s, a.y / s, a.z / s };
}
RG_HD float dot( v3 a, v3 b )
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
RG_HD v3 cross( v3 a, v3 b )
{
    return v3{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
RG_HD float length( v3 a )
{
    return sqrtf( dot( a, a ) );
}
RG_HD v3 normalize( v3 a )
{
    return a / length( a );
}
RG_HD float get( v3 a, int i )
{
    return i == 0 ? a.x : ( i == 1 ? a.y : a.z );
}
RG_HD v3 vmin( v3 a, v3 b )
{
    return v3{ fminf( a.x, b.x ), fminf( a.y, b.y ), fminf( a.z, b.z ) };
}
RG_HD v3 vmax( v3 a, v3 b )
{
    return v3{ fmaxf( a.x, b.x ), fmaxf( a.y, b.y ), fmaxf( a.z, b.z ) };
}
RG_HD float clampf( float x, float lo, float hi )
{
    return fminf( fmaxf( x, lo ), hi );
}
RG_HD float saturate( float x )
{
    return clampf( x, 0.0f, 1.0f );
}
RG_HD float mixf( float a, float b, float t )
{
    // GLSL mix(x, y, a) = x * (1 - a) + y * a
    return a * ( 1.0f - t ) + b * t;
}
RG_HD v3 mix3( v3 a, v3 b, float t )
{
    return v3{ mixf( a.x, b.x, t ), mixf( a.y, b.y, t ), mixf( a.z, b.z, t ) };
}
RG_HD float safePositiveRcp( float f )
{
    return f <= 0.0f ? 0.0f : 1.0f / f;
}
RG_HD float square( float x )
{
    return x * x;
}

// Luminance weights of Source/Shaders/Utils.h:54-57.
RG_HD float luminance( v3 c )
{
    return 0.2125f * c.x + 0.7154f * c.y + 0.0721f * c.z;
}

// ---------------------------------------------------------------- RNG
// Integer hashes of Source/Shaders/Random.h:243-300 (published by
// C. Wellons and A. Appleby; constants are the ones those publications give).
RG_HD uint32_t wellonsLowBias32( uint32_t x )
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// float( x ) / float( D ) for an integer 0 <= x <= D, without a divide: the
// product with the rounded reciprocal and one FMA correction step.  For
// D = 255 and D = 65535 it equals the correctly rounded quotient for every x
// (an exhaustive check over both ranges is part of the CPU tests), so the
// random numbers stay bit-exact while costing three instructions, not ten.
template< uint32_t D >
RG_HD float unitQuotient( uint32_t x )
{
    constexpr float C  = 1.0f / float( D );
    const float     xf = float( x );
    const float     q  = xf * C;
    const float     r  = fmaf( -q, float( D ), xf );
    return fmaf( r, C, q );
}

RG_HD float rnd16( uint32_t seed, uint32_t salt )
{
    uint32_t r = wellonsLowBias32( seed + salt );
    return unitQuotient< 65535u >( r & 0xFFFFu );
}

RG_HD void rnd16_2( uint32_t seed, uint32_t salt, float* a, float* b )
{
    uint32_t r = wellonsLowBias32( seed + salt );
    *a         = unitQuotient< 65535u >( r & 0xFFFFu );
    *b         = unitQuotient< 65535u >( r >> 16 );
}

RG_HD void rnd8_4( uint32_t seed, uint32_t salt, float out[ 4 ] )
{
    uint32_t r = wellonsLowBias32( seed + salt );
    out[ 0 ]   = unitQuotient< 255u >( r & 0xFFu );
    out[ 1 ]   = unitQuotient< 255u >( ( r >> 8 ) & 0xFFu );
    out[ 2 ]   = unitQuotient< 255u >( ( r >> 16 ) & 0xFFu );
    out[ 3 ]   = unitQuotient< 255u >( r >> 24 );
}

RG_HD void murmurHash33( uint32_t sx, uint32_t sy, uint32_t sz, uint32_t out[ 3 ] )
{
    const uint32_t M = 0x5bd1e995u;
    uint32_t       s[ 3 ] = { sx, sy, sz };
    for( int i = 0; i < 3; i++ )
    {
        s[ i ] *= M;
        s[ i ] ^= s[ i ] >> 24u;
        s[ i ] *= M;
    }
    uint32_t h[ 3 ] = { 1190494759u, 2147483647u, 3559788179u };
    for( int i = 0; i < 3; i++ )
    {
        h[ i ] *= M;
        h[ i ] ^= s[ 0 ];
        h[ i ] *= M;
        h[ i ] ^= s[ 1 ];
        h[ i ] *= M;
        h[ i ] ^= s[ 2 ];
        h[ i ] ^= h[ i ] >> 13u;
        h[ i ] *= M;
        h[ i ] ^= h[ i ] >> 15u;
        out[ i ] = h[ i ];
    }
}

// Blue-noise table coordinates packed as (layer << 14) | (y << 7) | x
// (Random.h:205-220, getRandomSeed at :290-300; 128 layers of 128 x 128).
RG_HD uint32_t randomSeed( int px, int py, uint32_t frameIndex )
{
    uint32_t h[ 3 ];
    murmurHash33( uint32_t( px ), uint32_t( py ), frameIndex, h );
    uint32_t ox = h[ 0 ] % 128u, oy = h[ 1 ] % 128u, layer = h[ 2 ] % 128u;
    return ( layer << 14 ) | ( oy << 7 ) | ox;
}

// ---------------------------------------------------------- checkerboard
// Source/Shaders/ShaderCommonGLSLFunc.h:306-366: odd pixels (x + y%2 odd)
// go to the right half of the image.
RG_HD int cbSeparator( int width )
{
    return width / 2;
}
RG_HD void cbFromRegular( int x, int y, int width, int* cx, int* cy )
{
    int odd = ( x + y % 2 ) % 2;
    *cx     = odd * cbSeparator( width ) + x / 2;
    *cy     = y;
}
RG_HD void cbToRegular( int cx, int cy, int width, int* x, int* y )
{
    int sep = cbSeparator( width );
    int odd = cx >= sep ? 1 : 0;
    int xx  = cx - odd * sep;
    *x      = xx * 2 + ( odd + cy ) % 2;
    *y      = cy;
}
// Render area of the checkerboard half containing cx: [lo, hi) x [0, height).
RG_HD void cbRenderArea( int cx, int width, int* lo, int* hi )
{
    int sep = cbSeparator( width );
    int odd = cx >= sep ? 1 : 0;
    *lo     = odd * sep;
    *hi     = ( odd + 1 ) * sep;
}
RG_HD bool inRenderArea( int x, int y, int lo, int hi, int height )
{
    return y >= 0 && y < height && x >= lo && x < hi;
}

} // namespace rg

namespace rg
{
// ------------------------------------------------- portable log2 / exp2
// FP32 log2 and exp2 evaluated with a fixed sequence of IEEE operations, so
// the GPU kernels and a scalar CPU evaluation produce the same bits (device
// and host libm differ in the last ulp, which would move histogram bin
// edges).  Accuracy is within a few ulp of the true value.
//
// log2: x = m * 2^e with m in [sqrt(1/2), sqrt(2)); ln(m) = 2 atanh(t) with
// t = (m - 1) / (m + 1), |t| < 0.1716, summed as t * (1 + t^2/3 + t^4/5 +
// ... + t^10/11) (the Taylor series of atanh; next term < 1e-10).
// Valid for finite x > 0 (normal or subnormal).
RG_HD float rgLog2( float x )
{
    uint32_t bits;
    std::memcpy( &bits, &x, 4 );
    int e = int( ( bits >> 23 ) & 0xFFu );
    if( e == 0 )
    {
        // subnormal: scale by 2^32 first
        float xs = x * 4294967296.0f;
        std::memcpy( &bits, &xs, 4 );
        e = int( ( bits >> 23 ) & 0xFFu ) - 32;
    }
    e -= 127;
    uint32_t mb = ( bits & 0x007FFFFFu ) | 0x3F800000u;
    float    m;
    std::memcpy( &m, &mb, 4 );
    if( m > 1.41421356f )
    {
        m = m * 0.5f;
        e += 1;
    }
    const float t  = ( m - 1.0f ) / ( m + 1.0f );
    const float t2 = t * t;
    const float s  = t2 * ( 0.333333333f + t2 * ( 0.2f + t2 * ( 0.142857143f + t2 * ( 0.111111111f + t2 * 0.0909090909f ) ) ) );
    // 2 / ln(2) = 2.88539008...
    return float( e ) + 2.88539008f * ( t + t * s );
}

// exp2: x = n + f with n = round(x), |f| <= 1/2; 2^f = e^(f ln 2) summed by
// its Taylor series to degree 8 (remainder < 2e-10 relative); 2^n is
// applied exactly through the exponent field.  Returns 0 below 2^-126 and
// +inf above 2^128.
RG_HD float rgExp2( float x )
{
    if( !( x > -126.0f ) )
    {
        return x != x ? x : 0.0f;
    }
    if( x >= 128.0f )
    {
        return INFINITY;
    }
    const float n = floorf( x + 0.5f );
    const float f = x - n;
    const float y = f * 0.693147181f;
    const float p = 1.0f + y * ( 1.0f + y * ( 0.5f + y * ( 0.166666667f + y * ( 0.0416666667f + y * ( 0.00833333333f + y * ( 0.00138888889f + y * ( 0.000198412698f + y * 0.0000248015873f ) ) ) ) ) ) );
    int ni = int( n );
    // split 2^n so that both factors are normal
    const int      h  = ni / 2;
    const uint32_t b1 = uint32_t( h + 127 ) << 23;
    const uint32_t b2 = uint32_t( ni - h + 127 ) << 23;
    float          s1, s2;
    std::memcpy( &s1, &b1, 4 );
    std::memcpy( &s2, &b2, 4 );
    return ( p * s1 ) * s2;
}
} // namespace rg
