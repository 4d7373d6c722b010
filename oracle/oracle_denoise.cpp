// CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// A-SVGF restated in FP64 from the reference compute shaders:
//   CmASVGFGradientAtrous.comp, CmSVGFTemporalAccumulation.comp,
//   CmAntiFirefly.comp, CmSVGFEstimateVariance.comp, CmSVGFAtrous_Iter0.comp,
//   CmSVGFAtrous.comp; pass order of Denoiser::Denoise (Denoiser.cpp:64-255).
// Every plane is a double array; a fetch outside the image returns zeros
// (texelFetch with robust image access).
#include "oracle.h"
#include "oracle_denoise.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace
{

// Rows of a pass on g_threads host threads (or_denoise_set_threads).  Every
// pass writes planes (or scratch arrays) that it does not read, so the rows
// are independent and the result does not depend on the thread count.
int g_threads = 1;
template< typename F >
void ParallelRows( int n, F&& f )
{
    const int threads = std::max( 1, std::min( g_threads, n ) );
    if( threads == 1 )
    {
        for( int i = 0; i < n; i++ )
        {
            f( i );
        }
        return;
    }
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

const double PI            = 3.14159265358979323846;
const double MAX_RAY_LEN   = 10000.0; // MAX_RAY_LENGTH, ShaderCommonC.h:197
const int    STRATA        = 3;       // COMPUTE_ASVGF_STRATA_SIZE

struct D3
{
    double x, y, z;
};
D3     operator+( D3 a, D3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
D3     operator-( D3 a, D3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
D3     operator*( D3 a, double s ) { return { a.x * s, a.y * s, a.z * s }; }
D3     operator*( D3 a, D3 b ) { return { a.x * b.x, a.y * b.y, a.z * b.z }; }
double dot( D3 a, D3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3     normalize( D3 a ) { return a * ( 1.0 / std::sqrt( dot( a, a ) ) ); }
double clampd( double x, double lo, double hi ) { return std::min( std::max( x, lo ), hi ); }
double saturate( double x ) { return clampd( x, 0.0, 1.0 ); }
double mix( double a, double b, double t ) { return a * ( 1.0 - t ) + b * t; }
D3     mix( D3 a, D3 b, double t ) { return { mix( a.x, b.x, t ), mix( a.y, b.y, t ), mix( a.z, b.z, t ) }; }
double lum( D3 c ) { return 0.2125 * c.x + 0.7154 * c.y + 0.0721 * c.z; } // Utils.h:54-57
double safeRcp( double f ) { return f <= 0.0 ? 0.0 : 1.0 / f; }
double smoothstep( double e0, double e1, double x )
{
    double t = clampd( ( x - e0 ) / ( e1 - e0 ), 0.0, 1.0 );
    return t * t * ( 3.0 - 2.0 * t );
}

D3 DecodeNormal( double packed )
{
    double n[ 3 ];
    or_decode_normal( uint32_t( packed ), n );
    return { n[ 0 ], n[ 1 ], n[ 2 ] };
}

// Column-major 4x4 times (v, w).
void MulCM( const float* m, D3 v, double w, double out[ 4 ] )
{
    for( int r = 0; r < 4; r++ )
    {
        out[ r ] = double( m[ 0 * 4 + r ] ) * v.x + double( m[ 1 * 4 + r ] ) * v.y + double( m[ 2 * 4 + r ] ) * v.z + double( m[ 3 * 4 + r ] ) * w;
    }
}

struct Fb
{
    const OrFramebuffers* f;
    int                   W, H, GW, GH;

    explicit Fb( const OrFramebuffers* fb ) : f( fb ), W( fb->width ), H( fb->height )
    {
        GW = std::max( 1, ( W + 1 ) / 3 );
        GH = std::max( 1, ( H + 1 ) / 3 );
    }
    int  comps( int p ) const { return or_plane_components( p ); }
    int  pw( int p ) const { return or_plane_layout( p ) == 2 ? GW : W; }
    int  ph( int p ) const { return or_plane_layout( p ) == 2 ? GH : H; }
    bool inside( int p, int x, int y ) const { return x >= 0 && y >= 0 && x < pw( p ) && y < ph( p ); }
    double get( int p, int x, int y, int c, bool prev = false ) const
    {
        if( !inside( p, x, y ) )
        {
            return 0.0;
        }
        const double* a = prev ? f->prev[ p ] : f->cur[ p ];
        return a[ ( size_t( y ) * pw( p ) + x ) * comps( p ) + c ];
    }
    D3 get3( int p, int x, int y, bool prev = false ) const
    {
        return { get( p, x, y, 0, prev ), get( p, x, y, 1, prev ), get( p, x, y, 2, prev ) };
    }
    void set( int p, int x, int y, int c, double v ) const
    {
        if( inside( p, x, y ) )
        {
            f->cur[ p ][ ( size_t( y ) * pw( p ) + x ) * comps( p ) + c ] = v;
        }
    }
    void set3( int p, int x, int y, D3 v ) const
    {
        set( p, x, y, 0, v.x );
        set( p, x, y, 1, v.y );
        set( p, x, y, 2, v.z );
    }
    void set4( int p, int x, int y, D3 v, double w ) const
    {
        set3( p, x, y, v );
        set( p, x, y, 3, w );
    }
};

// ShaderCommonGLSLFunc.h:346-366
void RenderArea( int cx, int W, int* lo, int* hi )
{
    const int sep = W / 2;
    const int odd = cx >= sep ? 1 : 0;
    *lo           = odd * sep;
    *hi           = ( odd + 1 ) * sep;
}
bool InArea( int x, int y, int lo, int hi, int H )
{
    return y >= 0 && y < H && x >= lo && x < hi;
}
// ShaderCommonGLSLFunc.h:368-376
bool TestDepth( double z, double zPrev, double zMotion )
{
    return std::fabs( z - zPrev + zMotion ) / std::fabs( z ) < 0.1;
}
bool TestNormal( D3 n, D3 nPrev )
{
    return dot( n, nPrev ) > 0.95;
}

// getPrevScreenPos, ShaderCommonGLSLFunc.h:290-297
void PrevScreenPos( double mx, double my, int x, int y, int W, int H, double* px, double* py )
{
    const double sx = double( W ) / 2.0, sy = double( H );
    *px             = ( ( x + 0.5 ) * ( 1.0 / sx ) + mx ) * sx;
    *py             = ( ( y + 0.5 ) * ( 1.0 / sy ) + my ) * sy;
}

// BRDF.h:27-29, :49-53, :57-64
double RoughnessSquaredToSpecPower( double a )
{
    return std::max( 0.01, 2.0 / ( a * a + 1e-4 ) - 2.0 );
}
D3 SpecularColor( D3 albedo, double metallic )
{
    return mix( D3{ 0.04, 0.04, 0.04 }, albedo, metallic );
}
double MaterialAmbient( D3 albedo )
{
    const double l = lum( albedo );
    return l > 0.02 ? 1.0 : 1.0 - ( ( l - 0.02 ) / 0.02 ) * ( ( l - 0.02 ) / 0.02 );
}

// CmSVGFTemporalAccumulation.comp:40-86
double Parallax( const RgExtFrameParams* g, D3 pos, D3 posPrev )
{
    D3     camDelta = D3{ g->cameraPositionPrev[ 0 ], g->cameraPositionPrev[ 1 ], g->cameraPositionPrev[ 2 ] } -
                  D3{ g->cameraPosition[ 0 ], g->cameraPosition[ 1 ], g->cameraPosition[ 2 ] };
    D3     v        = normalize( pos );
    D3     vPrev    = normalize( posPrev - camDelta );
    double cosa     = saturate( dot( v, vPrev ) );
    double parallax = std::sqrt( 1.0 - cosa * cosa ) / std::max( cosa, 1e-6 );
    return parallax * 60.0;
}
double SpecMinAccumFrames( const RgExtFrameParams* g )
{
    const double fps = safeRcp( g->timeDelta );
    return clampd( 0.2 * fps, 0.0, 128.0 );
}
double SpecAccumSpeed( const RgExtFrameParams* g, double Amax, double roughness, double NoV, double parallax )
{
    const double acos01sq            = 1.0 - NoV;
    const double a                   = std::pow( saturate( acos01sq ), 0.66 );
    const double b                   = 1.1 + roughness * roughness;
    const double parallaxSensitivity = ( b + a ) / ( b - a );
    const double powerScale          = 1.0 + parallax * parallaxSensitivity;
    double       f                   = 1.0 - std::exp2( -200.0 * roughness * roughness );
    f *= std::pow( saturate( roughness ), 0.5 * powerScale );
    const double A = 256.0 * f;
    return clampd( A, SpecMinAccumFrames( g ), Amax );
}
double SpecularDominantFactor( double nv, double roughness )
{
    const double a = 0.298475 * std::log( 39.4115 - 39.0029 * roughness );
    const double f = std::pow( clampd( 1.0 - nv, 0.0, 1.0 ), 10.8649 ) * ( 1.0 - a ) + a;
    return clampd( f, 0.0, 1.0 );
}

} // namespace

extern "C" int or_plane_components( int plane )
{
    switch( plane )
    {
        case OR_PLANE_GI_TIE:
        case RG_EXT_PLANE_NORMAL:
        case RG_EXT_PLANE_DEPTH_WORLD:
        case RG_EXT_PLANE_DEPTH_GRAD:
        case RG_EXT_PLANE_IS_SKY:
        case RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE:
        case RG_EXT_PLANE_HISTOGRAM_INPUT:
        case RG_EXT_PLANE_GRADIENT_PREV_PIX: return 1;
        case RG_EXT_PLANE_METALLIC_ROUGHNESS:
        case RG_EXT_PLANE_DIFF_ACCUM_MOMENTS:
        case RG_EXT_PLANE_GRADIENT_INPUTS:
        case RG_EXT_PLANE_INDIR_INITIAL_RADIANCE: return 2;
        case RG_EXT_PLANE_GI_REUSE: return 16;
        default: return 4;
    }
}

// 0 = checkerboard W x H, 1 = regular W x H, 2 = gradient strata.
extern "C" int or_plane_layout( int plane )
{
    switch( plane )
    {
        case RG_EXT_PLANE_ALBEDO:
        case RG_EXT_PLANE_HISTOGRAM_INPUT:
        case RG_EXT_PLANE_PRE_FINAL:
        case RG_EXT_PLANE_FINAL:
        case RG_EXT_PLANE_SCREEN_EMISSION:
        case RG_EXT_PLANE_ACID_FOG: return 1;
        case RG_EXT_PLANE_DIS_PING_GRADIENT:
        case RG_EXT_PLANE_DIS_PONG_GRADIENT:
        case RG_EXT_PLANE_DIS_GRADIENT_HISTORY:
        case RG_EXT_PLANE_GRADIENT_PREV_PIX: return 2;
        default: return 0;
    }
}

extern "C" int or_plane_has_prev( int plane )
{
    switch( plane )
    {
        case OR_PLANE_GI_TIE:
        case RG_EXT_PLANE_NORMAL:
        case RG_EXT_PLANE_METALLIC_ROUGHNESS:
        case RG_EXT_PLANE_DEPTH_WORLD:
        case RG_EXT_PLANE_SURFACE_POSITION:
        case RG_EXT_PLANE_VISIBILITY_BUFFER:
        case RG_EXT_PLANE_ACCUM_HISTORY_LENGTH:
        case RG_EXT_PLANE_DIFF_ACCUM_MOMENTS:
        case RG_EXT_PLANE_SPEC_ACCUM_COLOR:
        case RG_EXT_PLANE_INDIR_ACCUM:
        case RG_EXT_PLANE_RESERVOIRS:
        case RG_EXT_PLANE_GRADIENT_INPUTS:
        case RG_EXT_PLANE_INDIR_RESERVOIR_SAMPLE:
        case RG_EXT_PLANE_INDIR_RESERVOIR_WEIGHT: return 1;
        default: return 0;
    }
}

// CmASVGFGradientAtrous.comp:42-105
extern "C" void or_denoise_gradient_atrous( const OrFramebuffers* fbp, const RgExtFrameParams* g, uint32_t iter )
{
    Fb        fb( fbp );
    const int src  = ( iter % 2 == 0 ) ? RG_EXT_PLANE_DIS_PING_GRADIENT : RG_EXT_PLANE_DIS_PONG_GRADIENT;
    const int dst  = ( iter % 2 == 0 ) ? RG_EXT_PLANE_DIS_PONG_GRADIENT : RG_EXT_PLANE_DIS_PING_GRADIENT;
    const int step = 1 << iter;
    const double kernel[ 2 ][ 2 ] = { { 1.0, 0.5 }, { 0.5, 0.25 } };
    std::vector< double > out( size_t( fb.GW ) * fb.GH * 4 );
    ParallelRows( fb.GH, [ & ]( int gy ) {
        for( int gx = 0; gx < fb.GW; gx++ )
        {
            int lo, hi;
            RenderArea( gx * STRATA, fb.W, &lo, &hi );
            lo /= STRATA;
            hi /= STRATA;
            const int gh = fb.H / STRATA;
            D3        acc{ 0, 0, 0 };
            double    wsum = 0.0;
            for( int yy = -1; yy <= 1; yy++ )
            {
                for( int xx = -1; xx <= 1; xx++ )
                {
                    const int    qx = gx + xx * step, qy = gy + yy * step;
                    const double w  = kernel[ std::abs( yy ) ][ std::abs( xx ) ] * ( InArea( qx, qy, lo, hi, gh ) ? 1.0 : 0.0 );
                    acc             = acc + fb.get3( src, qx, qy ) * w;
                    wsum += w;
                }
            }
            D3 r = acc * ( 1.0 / wsum );
            double* o = &out[ ( size_t( gy ) * fb.GW + gx ) * 4 ];
            o[ 0 ] = r.x; o[ 1 ] = r.y; o[ 2 ] = r.z; o[ 3 ] = 0.0;
        }
    } );
    for( int gy = 0; gy < fb.GH; gy++ )
    {
        for( int gx = 0; gx < fb.GW; gx++ )
        {
            const double* o = &out[ ( size_t( gy ) * fb.GW + gx ) * 4 ];
            fb.set4( dst, gx, gy, { o[ 0 ], o[ 1 ], o[ 2 ] }, 0.0 );
            if( iter == 0 )
            {
                fb.set4( RG_EXT_PLANE_DIS_GRADIENT_HISTORY, gx, gy, { o[ 0 ], o[ 1 ], o[ 2 ] }, 0.0 );
            }
        }
    }
}

// CmSVGFTemporalAccumulation.comp:101-385
extern "C" void or_denoise_temporal( const OrFramebuffers* fbp, const RgExtFrameParams* g )
{
    Fb        fb( fbp );
    const int W = fb.W, H = fb.H;
    ParallelRows( H, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const D3     unfDiff  = fb.get3( RG_EXT_PLANE_UNFILTERED_DIRECT, x, y );
            const D3     unfIndir = fb.get3( RG_EXT_PLANE_UNFILTERED_INDIR, x, y );
            const D3     unfSpec  = fb.get3( RG_EXT_PLANE_UNFILTERED_SPECULAR, x, y );
            const D3     rayDir   = fb.get3( RG_EXT_PLANE_VIEW_DIRECTION, x, y );
            const double rayHitD  = fb.get( RG_EXT_PLANE_VIEW_DIRECTION, x, y, 3 );
            int          alo, ahi;
            RenderArea( x, W, &alo, &ahi );
            const double motionZ   = fb.get( RG_EXT_PLANE_MOTION, x, y, 2 );
            const double depth     = fb.get( RG_EXT_PLANE_DEPTH_WORLD, x, y, 0 );
            const D3     normal    = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, x, y, 0 ) );
            const double roughness = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 1 );
            const double l         = lum( unfDiff );
            const double mom[ 2 ]  = { l, l * l };

            double px, py;
            PrevScreenPos( fb.get( RG_EXT_PLANE_MOTION, x, y, 0 ), fb.get( RG_EXT_PLANE_MOTION, x, y, 1 ), x, y, W, H, &px, &py );
            const int    ppx = int( std::floor( px - 0.5 ) ), ppy = int( std::floor( py - 0.5 ) );
            const double sx = ( px - 0.5 ) - std::floor( px - 0.5 ), sy = ( py - 0.5 ) - std::floor( py - 0.5 );
            const double bw[ 2 ][ 2 ] = { { ( 1 - sx ) * ( 1 - sy ), sx * ( 1 - sy ) }, { ( 1 - sx ) * sy, sx * sy } };

            double hist[ 4 ] = { 0, 0, 0, 0 };
            D3     diffPrev{ 0, 0, 0 }, specPrevOnSurf{ 0, 0, 0 }, indirPrev{ 0, 0, 0 };
            double momPrev[ 2 ] = { 0, 0 };
            double weightSum    = 0.0;
            for( int yy = 0; yy <= 1; yy++ )
            {
                for( int xx = 0; xx <= 1; xx++ )
                {
                    const int qx = ppx + xx, qy = ppy + yy;
                    const double dPrev = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0, true );
                    const D3     nPrev = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0, true ) );
                    if( InArea( qx, qy, alo, ahi, H ) && TestDepth( depth, dPrev, motionZ ) && TestNormal( normal, nPrev ) )
                    {
                        const double w = bw[ yy ][ xx ];
                        diffPrev       = diffPrev + fb.get3( RG_EXT_PLANE_DIFF_COLOR_HISTORY, qx, qy ) * w;
                        momPrev[ 0 ] += w * fb.get( RG_EXT_PLANE_DIFF_ACCUM_MOMENTS, qx, qy, 0, true );
                        momPrev[ 1 ] += w * fb.get( RG_EXT_PLANE_DIFF_ACCUM_MOMENTS, qx, qy, 1, true );
                        for( int c = 0; c < 4; c++ )
                        {
                            hist[ c ] += w * fb.get( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, qx, qy, c, true );
                        }
                        specPrevOnSurf = specPrevOnSurf + fb.get3( RG_EXT_PLANE_SPEC_ACCUM_COLOR, qx, qy, true ) * w;
                        indirPrev      = indirPrev + fb.get3( RG_EXT_PLANE_INDIR_ACCUM, qx, qy, true ) * w;
                        weightSum += w;
                    }
                }
            }

            const double aDiff  = fb.get( RG_EXT_PLANE_DIS_PING_GRADIENT, x / STRATA, y / STRATA, 0 );
            const double aIndir = fb.get( RG_EXT_PLANE_DIS_PING_GRADIENT, x / STRATA, y / STRATA, 1 );
            const double aSpec  = fb.get( RG_EXT_PLANE_DIS_PING_GRADIENT, x / STRATA, y / STRATA, 2 );

            D3     specAccum;
            double specHistAccum, specHitAccum;
            if( weightSum > 0.0001 )
            {
                specPrevOnSurf            = specPrevOnSurf * ( 1.0 / weightSum );
                const double specHistLen  = hist[ 2 ] / weightSum;
                const double specHitPrev  = hist[ 3 ] / weightSum;
                const D3     surfPos      = fb.get3( RG_EXT_PLANE_SURFACE_POSITION, x, y );
                const D3     surfPosPrev  = fb.get3( RG_EXT_PLANE_SURFACE_POSITION, ppx, ppy, true );
                const double nv           = std::max( 0.0, dot( normal, rayDir * -1.0 ) );
                const double A            = clampd( 1.0 + specHistLen * std::pow( 1.0 - aSpec, 5.0 ), SpecMinAccumFrames( g ), 256.0 );
                const double parallax     = Parallax( g, surfPos, surfPosPrev );
                const double AOnSurf      = SpecAccumSpeed( g, A, roughness, nv, parallax );
                const D3     specOnSurf   = mix( specPrevOnSurf, unfSpec, 1.0 / ( 1.0 + AOnSurf ) );
                const double hitOnSurf    = mix( specHitPrev, rayHitD, 1.0 / ( 1.0 + AOnSurf ) );
                specAccum                 = specOnSurf;
                specHistAccum             = AOnSurf;
                specHitAccum              = hitOnSurf;
                const D3 virtualPos       = surfPos + rayDir * ( hitOnSurf * SpecularDominantFactor( nv, roughness ) );

                double vc[ 4 ], vp[ 4 ], cc[ 4 ], cp[ 4 ];
                MulCM( g->view, virtualPos, 1.0, vc );
                MulCM( g->viewPrev, virtualPos, 1.0, vp );
                MulCM( g->projection, { vc[ 0 ], vc[ 1 ], vc[ 2 ] }, vc[ 3 ], cc );
                MulCM( g->projectionPrev, { vp[ 0 ], vp[ 1 ], vp[ 2 ] }, vp[ 3 ], cp );
                const double scx = ( cc[ 0 ] / cc[ 3 ] ) * 0.5 + 0.5, scy = ( cc[ 1 ] / cc[ 3 ] ) * 0.5 + 0.5;
                const double spx = ( cp[ 0 ] / cp[ 3 ] ) * 0.5 + 0.5, spy = ( cp[ 1 ] / cp[ 3 ] ) * 0.5 + 0.5;
                double       qpx, qpy;
                PrevScreenPos( spx - scx, spy - scy, x, y, W, H, &qpx, &qpy );
                const int    vpx = int( std::floor( qpx - 0.5 ) ), vpy = int( std::floor( qpy - 0.5 ) );
                const double tx = ( qpx - 0.5 ) - std::floor( qpx - 0.5 ), ty = ( qpy - 0.5 ) - std::floor( qpy - 0.5 );
                const double bws[ 2 ][ 2 ] = { { ( 1 - tx ) * ( 1 - ty ), tx * ( 1 - ty ) }, { ( 1 - tx ) * ty, tx * ty } };
                D3     specPrevVirtual{ 0, 0, 0 };
                double wsv = 0.0;
                for( int yy = 0; yy <= 1; yy++ )
                {
                    for( int xx = 0; xx <= 1; xx++ )
                    {
                        const int    qx = vpx + xx, qy = vpy + yy;
                        const double dPrev = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0, true );
                        const D3     nPrev = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0, true ) );
                        if( InArea( qx, qy, alo, ahi, H ) && TestDepth( depth, dPrev, motionZ ) && TestNormal( normal, nPrev ) )
                        {
                            specPrevVirtual = specPrevVirtual + fb.get3( RG_EXT_PLANE_SPEC_ACCUM_COLOR, qx, qy, true ) * bws[ yy ][ xx ];
                            wsv += bws[ yy ][ xx ];
                        }
                    }
                }
                if( wsv > 0.0001 )
                {
                    ( void )specPrevVirtual; // the reference mixes specPrevOnSurf below, not the virtual sample
                    double       Avirt      = SpecAccumSpeed( g, A, roughness, nv, 0.0 );
                    const double confidence = 1.0;
                    const double Amin       = std::min( Avirt, 3.0 * std::sqrt( roughness ) );
                    double       a          = mix( 1.0 / ( 1.0 + Amin ), 1.0 / ( 1.0 + Avirt ), confidence );
                    Avirt                   = 1.0 / a - 1.0;
                    const double Ahit       = std::min( Avirt, 16.0 );
                    const D3     specVirt   = mix( specPrevOnSurf, unfSpec, 1.0 / ( 1.0 + Avirt ) );
                    const double amount     = SpecularDominantFactor( nv, roughness );
                    a                       = mix( 1.0 / ( 1.0 + AOnSurf ), 1.0 / ( 1.0 + Avirt ), amount );
                    specAccum               = mix( specOnSurf, specVirt, amount );
                    specHistAccum           = 1.0 / a - 1.0;
                    specHitAccum            = mix( specHitPrev, rayHitD, 1.0 / ( 1.0 + Ahit ) );
                }
            }
            else
            {
                specAccum     = unfSpec;
                specHistAccum = 1.0;
                specHitAccum  = rayHitD;
            }

            D3     diffAccum, indirAccum;
            double diffMom[ 2 ], diffHistAccum, indirHistAccum;
            if( weightSum > 0.0001 )
            {
                const double inv = 1.0 / weightSum;
                diffPrev         = diffPrev * inv;
                momPrev[ 0 ] *= inv;
                momPrev[ 1 ] *= inv;
                indirPrev        = indirPrev * inv;
                double dh        = hist[ 0 ] * inv;
                double ih        = hist[ 1 ] * inv;

                dh *= std::pow( 1.0 - aDiff, 10.0 );
                dh = clampd( dh + 1.0, 1.0, 256.0 );
                double alphaC = std::max( 0.01, 1.0 / dh );
                double alphaM = std::max( 0.01, 1.0 / dh );
                alphaC        = mix( alphaC, 1.0, aDiff );
                alphaM        = mix( alphaM, 1.0, aDiff );
                diffAccum     = mix( diffPrev, unfDiff, alphaC );
                diffMom[ 0 ]  = mix( momPrev[ 0 ], mom[ 0 ], alphaM );
                diffMom[ 1 ]  = mix( momPrev[ 1 ], mom[ 1 ], alphaM );
                diffHistAccum = dh;

                ih *= std::pow( 1.0 - aIndir, 10.0 );
                ih = clampd( ih + 1.0, 1.0, 256.0 );
                double aI      = std::max( 0.01, 1.0 / ih );
                aI             = mix( aI, 1.0, aIndir );
                indirAccum     = mix( indirPrev, unfIndir, aI );
                indirHistAccum = ih;
            }
            else
            {
                diffAccum      = unfDiff;
                diffMom[ 0 ]   = mom[ 0 ];
                diffMom[ 1 ]   = mom[ 1 ];
                diffHistAccum  = 1.0;
                indirAccum     = unfIndir;
                indirHistAccum = 1.0;
            }
            fb.set4( RG_EXT_PLANE_DIFF_TEMPORARY, x, y, diffAccum, 0.0 );
            fb.set( RG_EXT_PLANE_DIFF_ACCUM_MOMENTS, x, y, 0, diffMom[ 0 ] );
            fb.set( RG_EXT_PLANE_DIFF_ACCUM_MOMENTS, x, y, 1, diffMom[ 1 ] );
            fb.set( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, x, y, 0, diffHistAccum );
            fb.set( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, x, y, 1, indirHistAccum );
            fb.set( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, x, y, 2, specHistAccum );
            fb.set( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, x, y, 3, specHitAccum );
            fb.set4( RG_EXT_PLANE_SPEC_ACCUM_COLOR, x, y, specAccum, 0.0 );
            fb.set4( RG_EXT_PLANE_INDIR_ACCUM, x, y, indirAccum, 0.0 );
        }
    } );
}

// CmAntiFirefly.comp:96-159 (5x5 min/max clamp, radius 2, centre excluded)
extern "C" void or_denoise_antifirefly( const OrFramebuffers* fbp, const RgExtFrameParams* g )
{
    Fb        fb( fbp );
    const int W = fb.W, H = fb.H;
    const int srcs[ 3 ] = { RG_EXT_PLANE_DIFF_TEMPORARY, RG_EXT_PLANE_SPEC_ACCUM_COLOR, RG_EXT_PLANE_INDIR_ACCUM };
    const int dsts[ 3 ] = { RG_EXT_PLANE_DIFF_ACCUM_COLOR, RG_EXT_PLANE_SPEC_PING, RG_EXT_PLANE_INDIR_PING };
    ParallelRows( H, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            for( int s = 0; s < 3; s++ )
            {
                D3 c = fb.get3( srcs[ s ], x, y );
                if( g->antiFireflyEnabled )
                {
                    D3 mn{ 1e6, 1e6, 1e6 }, mx{ -1e6, -1e6, -1e6 };
                    for( int i = -2; i <= 2; i++ )
                    {
                        for( int j = -2; j <= 2; j++ )
                        {
                            if( i == 0 && j == 0 )
                            {
                                continue;
                            }
                            // GET_SDATA(i, j): i is the x offset, j the y offset
                            const D3 o = fb.get3( srcs[ s ], x + i, y + j );
                            mn = { std::min( mn.x, o.x ), std::min( mn.y, o.y ), std::min( mn.z, o.z ) };
                            mx = { std::max( mx.x, o.x ), std::max( mx.y, o.y ), std::max( mx.z, o.z ) };
                        }
                    }
                    c = { clampd( c.x, mn.x, mx.x ), clampd( c.y, mn.y, mx.y ), clampd( c.z, mn.z, mx.z ) };
                }
                fb.set4( dsts[ s ], x, y, c, 0.0 );
            }
        }
    } );
}

// CmSVGFEstimateVariance.comp:129-206
extern "C" void or_denoise_variance( const OrFramebuffers* fbp, const RgExtFrameParams* g )
{
    Fb        fb( fbp );
    const int W = fb.W, H = fb.H;
    const double HT = 4.0; // HISTORY_LENGTH_THRESHOLD
    ParallelRows( H, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            D3           color = fb.get3( RG_EXT_PLANE_DIFF_ACCUM_COLOR, x, y );
            const double depth = fb.get( RG_EXT_PLANE_DEPTH_WORLD, x, y, 0 );
            if( depth < 0.0 || depth > MAX_RAY_LEN )
            {
                fb.set4( RG_EXT_PLANE_DIFF_PING, x, y, color, 0.0 );
                continue;
            }
            const double histLen = fb.get( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, x, y, 0 );
            const double m0 = fb.get( RG_EXT_PLANE_DIFF_ACCUM_MOMENTS, x, y, 0 ), m1 = fb.get( RG_EXT_PLANE_DIFF_ACCUM_MOMENTS, x, y, 1 );
            if( histLen > HT )
            {
                double var = std::max( 0.0, m1 - m0 * m0 );
                var /= std::max( 1.0, histLen - HT );
                fb.set4( RG_EXT_PLANE_DIFF_PING, x, y, color, var );
                continue;
            }
            const double l    = lum( color );
            double       s0   = l + m0, s1 = l * l + m1;
            double       wsum = 1.0;
            const D3     n    = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, x, y, 0 ) );
            for( int yy = -1; yy <= 1; yy++ )
            {
                for( int xx = -1; xx <= 1; xx++ )
                {
                    if( xx == 0 && yy == 0 )
                    {
                        continue;
                    }
                    const D3     cq  = fb.get3( RG_EXT_PLANE_DIFF_ACCUM_COLOR, x + xx, y + yy );
                    const double lq  = lum( cq );
                    const double dq  = fb.get( RG_EXT_PLANE_DEPTH_WORLD, x + xx, y + yy, 0 );
                    const D3     nq  = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, x + xx, y + yy, 0 ) );
                    const double w_z = std::exp( -std::fabs( depth - dq ) * safeRcp( std::max( depth, dq ) ) );
                    const double w_n = std::pow( std::max( 0.0, dot( n, nq ) ), 128.0 );
                    const double w   = w_z * w_n;
                    if( !std::isnan( w ) )
                    {
                        s0 += lq * w;
                        s1 += lq * lq * w;
                        color = color + cq * w;
                        wsum += w;
                    }
                }
            }
            const double inv = safeRcp( wsum );
            s0 *= inv;
            s1 *= inv;
            color = color * inv;
            const double var = std::max( 0.0, s1 - s0 * s0 ) * ( 1.0 + 2.0 * ( 1.0 - histLen / HT ) );
            fb.set4( RG_EXT_PLANE_DIFF_PING, x, y, color, var );
        }
    } );
}

namespace
{
// Source planes of a-trous iteration k (CmSVGFAtrous.comp:257-278 generalised
// so a fifth iteration continues the ping-pong).
void AtrousPlanes( uint32_t k, uint32_t count, int* srcD, int* srcS, int* srcI, int* dstD, int* dstS, int* dstI )
{
    if( k == 0 )
    {
        *srcD = RG_EXT_PLANE_DIFF_PING; *srcS = RG_EXT_PLANE_SPEC_PING; *srcI = RG_EXT_PLANE_INDIR_PING;
        *dstD = RG_EXT_PLANE_DIFF_COLOR_HISTORY; *dstS = RG_EXT_PLANE_SPEC_PONG; *dstI = RG_EXT_PLANE_INDIR_PONG;
        return;
    }
    if( k == 1 )
    {
        *srcD = RG_EXT_PLANE_DIFF_COLOR_HISTORY; *srcS = RG_EXT_PLANE_SPEC_PONG; *srcI = RG_EXT_PLANE_INDIR_PONG;
    }
    else if( k % 2 == 0 )
    {
        *srcD = RG_EXT_PLANE_DIFF_PING; *srcS = RG_EXT_PLANE_SPEC_PING; *srcI = RG_EXT_PLANE_INDIR_PING;
    }
    else
    {
        *srcD = RG_EXT_PLANE_DIFF_PONG; *srcS = RG_EXT_PLANE_SPEC_PONG; *srcI = RG_EXT_PLANE_INDIR_PONG;
    }
    if( k + 1 == count )
    {
        *dstD = *dstS = *dstI = -1; // last iteration composes
    }
    else if( k % 2 == 1 )
    {
        *dstD = RG_EXT_PLANE_DIFF_PING; *dstS = RG_EXT_PLANE_SPEC_PING; *dstI = RG_EXT_PLANE_INDIR_PING;
    }
    else
    {
        *dstD = RG_EXT_PLANE_DIFF_PONG; *dstS = RG_EXT_PLANE_SPEC_PONG; *dstI = RG_EXT_PLANE_INDIR_PONG;
    }
}
} // namespace

// CmSVGFAtrous_Iter0.comp:196-335 (k = 0) and CmSVGFAtrous.comp:71-349 (k >= 1).
extern "C" void or_denoise_atrous( const OrFramebuffers* fbp, const RgExtFrameParams* g, uint32_t k )
{
    Fb         fb( fbp );
    const int  W = fb.W, H = fb.H;
    const uint32_t count = g->atrousIterations;
    int srcD, srcS, srcI, dstD, dstS, dstI;
    AtrousPlanes( k, count, &srcD, &srcS, &srcI, &dstD, &dstS, &dstI );
    const int    step = 1 << k;
    const double kernel[ 2 ][ 2 ] = { { 1.0, 0.5 }, { 0.5, 0.25 } };
    const size_t np = size_t( W ) * H;
    std::vector< double > oD( np * 4 ), oS( np * 3 ), oI( np * 3 ), fvar( np, -1.0 );

    ParallelRows( H, [ & ]( int y ) {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            int          alo, ahi;
            RenderArea( x, W, &alo, &ahi );
            D3     outD = fb.get3( srcD, x, y );
            double outV = fb.get( srcD, x, y, 3 );
            D3     outS = fb.get3( srcS, x, y );
            D3     outI = fb.get3( srcI, x, y );
            const double depth = fb.get( RG_EXT_PLANE_DEPTH_WORLD, x, y, 0 );
            const double gradD = fb.get( RG_EXT_PLANE_DEPTH_GRAD, x, y, 0 );
            if( depth < 0.0 || depth > MAX_RAY_LEN )
            {
                if( k == 0 )
                {
                    outD = outS = outI = { 0, 0, 0 };
                    outV = 0.0;
                }
            }
            else
            {
                const D3     n     = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, x, y, 0 ) );
                const double l     = lum( outD );
                double       pvar;
                if( k == 0 )
                {
                    // prefilterLuminanceVariance: 3x3 Gaussian of the variance, sqrt
                    const double gk[ 2 ][ 2 ] = { { 0.25, 0.125 }, { 0.125, 0.0625 } };
                    double       r            = 0.0;
                    for( int yy = -1; yy <= 1; yy++ )
                    {
                        for( int xx = -1; xx <= 1; xx++ )
                        {
                            r += fb.get( srcD, x + xx, y + yy, 3 ) * gk[ std::abs( xx ) ][ std::abs( yy ) ];
                        }
                    }
                    pvar      = std::sqrt( std::max( r, 0.0 ) );
                    fvar[ i ] = pvar;
                }
                else
                {
                    pvar = fb.get( RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE, x, y, 0 );
                }
                const double wLum      = 1.0 / ( 4.0 * pvar + 0.00001 );
                const double roughness = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 1 );
                const double wRough    = clampd( roughness * 30.0 - double( k ), 0.0, 1.0 );
                const double histSpec  = fb.get( RG_EXT_PLANE_ACCUM_HISTORY_LENGTH, x, y, 2 );
                double       nws       = clampd( RoughnessSquaredToSpecPower( roughness * roughness ), 8.0, 1024.0 );
                nws *= clampd( histSpec / 8.0, 0.0, 1.0 );
                double ws = 1.0, wss = 1.0, wsi = 1.0;
                for( int yy = -1; yy <= 1; yy++ )
                {
                    for( int xx = -1; xx <= 1; xx++ )
                    {
                        if( xx == 0 && yy == 0 )
                        {
                            continue;
                        }
                        const int    qx = x + xx * step, qy = y + yy * step;
                        const double rq = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, qx, qy, 1 );
                        const D3     nq = DecodeNormal( fb.get( RG_EXT_PLANE_NORMAL, qx, qy, 0 ) );
                        const double dq = fb.get( RG_EXT_PLANE_DEPTH_WORLD, qx, qy, 0 );
                        const D3     sq = fb.get3( srcS, qx, qy );
                        const D3     cq = fb.get3( srcD, qx, qy );
                        const double vq = fb.get( srcD, qx, qy, 3 );
                        const double lq = lum( cq );
                        const double nn = std::max( 0.0, dot( n, nq ) );
                        const D3     iq = fb.get3( srcI, qx, qy );
                        const double w_z = std::fabs( depth - dq ) / std::max( gradD * ( std::abs( xx ) + std::abs( yy ) ), 0.01 );
                        const double w_n = std::pow( nn, 128.0 );
                        const double w_l = std::fabs( l - lq ) * wLum;
                        double       w_r = std::max( 0.0, 1.0 - 10.0 * std::fabs( roughness - rq ) ) * wRough;
                        if( nws > 0.0 )
                        {
                            w_r *= std::pow( nn, nws );
                        }
                        const double inside = InArea( qx, qy, alo, ahi, H ) ? 1.0 : 0.0;
                        const double wBase  = std::exp( -w_z * w_z ) * w_n * kernel[ std::abs( yy ) ][ std::abs( xx ) ] * inside;
                        const double wD = wBase * std::exp( -w_l ), wS = wBase * w_r, wI = wBase;
                        outD = outD + cq * wD;
                        outS = outS + sq * wS;
                        outI = outI + iq * wI;
                        outV += vq * wD * wD;
                        ws += wD;
                        wss += wS;
                        wsi += wI;
                    }
                }
                outD = outD * ( 1.0 / ws );
                outV *= ( 1.0 / ws ) * ( 1.0 / ws );
                outS = outS * ( 1.0 / wss );
                outI = outI * ( 1.0 / wsi );
            }
            oD[ i * 4 + 0 ] = outD.x; oD[ i * 4 + 1 ] = outD.y; oD[ i * 4 + 2 ] = outD.z; oD[ i * 4 + 3 ] = outV;
            oS[ i * 3 + 0 ] = outS.x; oS[ i * 3 + 1 ] = outS.y; oS[ i * 3 + 2 ] = outS.z;
            oI[ i * 3 + 0 ] = outI.x; oI[ i * 3 + 1 ] = outI.y; oI[ i * 3 + 2 ] = outI.z;
        }
    } );

    for( int y = 0; y < H; y++ )
    {
        for( int x = 0; x < W; x++ )
        {
            const size_t i = size_t( y ) * W + x;
            if( k == 0 && fvar[ i ] >= 0.0 )
            {
                fb.set( RG_EXT_PLANE_ATROUS_FILTERED_VARIANCE, x, y, 0, fvar[ i ] );
            }
            const D3 d{ oD[ i * 4 ], oD[ i * 4 + 1 ], oD[ i * 4 + 2 ] };
            const D3 s{ oS[ i * 3 ], oS[ i * 3 + 1 ], oS[ i * 3 + 2 ] };
            const D3 ind{ oI[ i * 3 ], oI[ i * 3 + 1 ], oI[ i * 3 + 2 ] };
            if( dstD >= 0 )
            {
                fb.set4( dstD, x, y, d, oD[ i * 4 + 3 ] );
                fb.set4( dstS, x, y, s, 0.0 );
                fb.set4( dstI, x, y, ind, 0.0 );
                continue;
            }
            // Composition (CmSVGFAtrous.comp:284-348), regular-layout outputs.
            int rx, ry;
            or_checkerboard_to_regular( x, y, W, &rx, &ry );
            const D3 albedo     = fb.get3( RG_EXT_PLANE_ALBEDO, rx, ry );
            const D3 throughput = fb.get3( RG_EXT_PLANE_THROUGHPUT, x, y );
            D3       ill;
            if( fb.get( RG_EXT_PLANE_IS_SKY, x, y, 0 ) == 0.0 )
            {
                const double metallic  = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 0 );
                const double roughness = fb.get( RG_EXT_PLANE_METALLIC_ROUGHNESS, x, y, 1 );
                const D3     spec      = mix( s, ind, smoothstep( 0.5, 0.75, roughness ) );
                const D3     ros       = SpecularColor( albedo, metallic );
                const D3     rod       = albedo * ( 1.0 - metallic );
                ill                    = ( d + ind ) * rod + spec * ros;
                ill                    = ill * throughput;
                ill                    = ill * MaterialAmbient( albedo );
            }
            else
            {
                ill = albedo * throughput;
            }
            fb.set( RG_EXT_PLANE_HISTOGRAM_INPUT, rx, ry, 0, std::max( 0.0, lum( ill ) ) );
            fb.set4( RG_EXT_PLANE_PRE_FINAL, rx, ry, { std::max( 0.0, ill.x ), std::max( 0.0, ill.y ), std::max( 0.0, ill.z ) }, 0.0 );
        }
    }
}

extern "C" void or_denoise( const OrFramebuffers* fb, const RgExtFrameParams* g )
{
    for( uint32_t i = 0; i < 4; i++ )
    {
        or_denoise_gradient_atrous( fb, g, i );
    }
    or_denoise_temporal( fb, g );
    or_denoise_antifirefly( fb, g );
    or_denoise_variance( fb, g );
    for( uint32_t k = 0; k < g->atrousIterations; k++ )
    {
        or_denoise_atrous( fb, g, k );
    }
}

// Host threads for the passes above (1 by default); the results do not
// depend on it.
extern "C" void or_denoise_set_threads( int threads )
{
    g_threads = std::max( 1, std::min( threads, 256 ) );
}
