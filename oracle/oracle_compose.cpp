// CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Final compose in FP64: CmCheckerboard.comp:49-76 (resolve of split pixels),
// CmPrepareFinal.comp:148-176 with auto exposure (Exposure.h:29-56) and the
// AMD FidelityFX LPM tone mapper as the reference vendors and sets it up
// (Shaders/LPM/ffx_lpm.h LpmSetup / LpmMap, ImageComposition.cpp:377-394:
// LPM_CONFIG_709_709, LPM_COLORS_709_709, softGap 0, hdrMax 256, exposure 8,
// contrast 0.1, shoulder contrast 1).
#include "oracle_compose.h"

#include <algorithm>
#include <cmath>

namespace
{
struct Lpm
{
    double sat[ 3 ], contrast, tsb[ 2 ], lumaT[ 3 ], rcpLumaT[ 3 ], crosstalk[ 3 ];
};

// LpmColRgbToXyz (ffx_lpm.h:195-205): the Y row of the RGB -> XYZ matrix
void RgbToXyzY( const double xy[ 4 ][ 2 ], double y[ 3 ] )
{
    double m[ 3 ][ 3 ];
    for( int c = 0; c < 3; c++ )
    {
        m[ 0 ][ c ] = xy[ c ][ 0 ];
        m[ 1 ][ c ] = xy[ c ][ 1 ];
        m[ 2 ][ c ] = 1.0 - xy[ c ][ 0 ] - xy[ c ][ 1 ];
    }
    const double w[ 3 ] = { xy[ 3 ][ 0 ] / xy[ 3 ][ 1 ], 1.0, ( 1.0 - xy[ 3 ][ 0 ] - xy[ 3 ][ 1 ] ) / xy[ 3 ][ 1 ] };
    // solve m * s = w (Cramer)
    auto det3 = []( double a[ 3 ][ 3 ] ) {
        return a[ 0 ][ 0 ] * ( a[ 1 ][ 1 ] * a[ 2 ][ 2 ] - a[ 1 ][ 2 ] * a[ 2 ][ 1 ] ) - a[ 0 ][ 1 ] * ( a[ 1 ][ 0 ] * a[ 2 ][ 2 ] - a[ 1 ][ 2 ] * a[ 2 ][ 0 ] ) +
               a[ 0 ][ 2 ] * ( a[ 1 ][ 0 ] * a[ 2 ][ 1 ] - a[ 1 ][ 1 ] * a[ 2 ][ 0 ] );
    };
    const double d = det3( m );
    for( int c = 0; c < 3; c++ )
    {
        double a[ 3 ][ 3 ];
        for( int r = 0; r < 3; r++ )
        {
            for( int k = 0; k < 3; k++ )
            {
                a[ r ][ k ] = k == c ? w[ r ] : m[ r ][ k ];
            }
        }
        y[ c ] = m[ 1 ][ c ] * det3( a ) / d;
    }
}

Lpm Setup( const float saturation[ 3 ], const float crosstalk[ 3 ] )
{
    Lpm          k;
    const double hdrMax = 256.0, exposure = 8.0, shoulderContrast = 1.0;
    const double contrast = 0.1 + 1.0; // "contrast needs to be 1.0 based"
    const double midIn = hdrMax * 0.18 * std::exp2( -exposure ), midOut = 0.18;
    const double cs = contrast * shoulderContrast;
    const double z0 = -std::pow( midIn, contrast );
    const double z1 = std::pow( hdrMax, cs ) * std::pow( midIn, contrast );
    const double z2 = std::pow( hdrMax, contrast ) * std::pow( midIn, cs ) * midOut;
    const double z3 = std::pow( hdrMax, cs ) * midOut;
    const double z4 = std::pow( midIn, cs ) * midOut;
    k.tsb[ 0 ]      = -( ( z0 + ( midOut * ( z1 - z2 ) ) / ( z3 - z4 ) ) / z4 );
    const double w0 = std::pow( hdrMax, cs ) * std::pow( midIn, contrast );
    const double w1 = std::pow( hdrMax, contrast ) * std::pow( midIn, cs ) * midOut;
    const double w2 = std::pow( hdrMax, cs ) * midOut;
    const double w3 = std::pow( midIn, cs ) * midOut;
    k.tsb[ 1 ]      = ( w0 - w1 ) / ( w2 - w3 );
    const double xy709[ 4 ][ 2 ] = { { 0.64, 0.33 }, { 0.30, 0.60 }, { 0.15, 0.06 }, { 0.3127, 0.3290 } }; // lpmCol709R/G/B, lpmColD65
    RgbToXyzY( xy709, k.lumaT );
    const double sum = k.lumaT[ 0 ] + k.lumaT[ 1 ] + k.lumaT[ 2 ];
    for( int c = 0; c < 3; c++ )
    {
        k.lumaT[ c ] /= sum;
        k.rcpLumaT[ c ]  = 1.0 / k.lumaT[ c ];
        k.sat[ c ]       = double( saturation[ c ] ) + contrast;
        k.crosstalk[ c ] = crosstalk[ c ];
    }
    k.contrast = contrast;
    return k;
}

// ASatF1 = clamp(x, 0, 1) with NaN -> 0, as the GPU clamp does: the LPM map
// reaches 0 * (1 / 0) when all three channels clip (grey above hdrMax), and
// this keeps such pixels white instead of propagating NaN
double sat( double x ) { return x > 0.0 ? std::min( x, 1.0 ) : 0.0; }

// LpmMap (ffx_lpm.h:726-828) for LPM_CONFIG_709_709 (no shoulder, con, soft, con2)
void Map( const Lpm& k, const double in[ 3 ], double out[ 3 ] )
{
    const double mx = std::max( in[ 0 ], std::max( in[ 1 ], in[ 2 ] ) );
    if( !( mx > 0.0 ) )
    {
        out[ 0 ] = out[ 1 ] = out[ 2 ] = 0.0;
        return;
    }
    double ratio[ 3 ];
    for( int c = 0; c < 3; c++ )
    {
        ratio[ c ] = std::pow( in[ c ] / mx, k.sat[ c ] );
    }
    double luma = in[ 0 ] * k.lumaT[ 0 ] + in[ 1 ] * k.lumaT[ 1 ] + in[ 2 ] * k.lumaT[ 2 ];
    luma        = std::pow( luma, k.contrast );
    luma        = luma / ( luma * k.tsb[ 0 ] + k.tsb[ 1 ] );
    const double lumaRatio  = ratio[ 0 ] * k.lumaT[ 0 ] + ratio[ 1 ] * k.lumaT[ 1 ] + ratio[ 2 ] * k.lumaT[ 2 ];
    const double ratioScale = sat( luma / lumaRatio );
    double       col[ 3 ], cap[ 3 ];
    for( int c = 0; c < 3; c++ )
    {
        col[ c ] = sat( ratio[ c ] * ratioScale );
        cap[ c ] = k.crosstalk[ c ] - k.crosstalk[ c ] * col[ c ];
    }
    double       lumaAdd = sat( luma - ( col[ 0 ] * k.lumaT[ 0 ] + col[ 1 ] * k.lumaT[ 1 ] + col[ 2 ] * k.lumaT[ 2 ] ) );
    const double t       = lumaAdd / ( cap[ 0 ] * k.lumaT[ 0 ] + cap[ 1 ] * k.lumaT[ 1 ] + cap[ 2 ] * k.lumaT[ 2 ] );
    for( int c = 0; c < 3; c++ )
    {
        col[ c ] = sat( t * cap[ c ] + col[ c ] );
    }
    lumaAdd = sat( luma - ( col[ 0 ] * k.lumaT[ 0 ] + col[ 1 ] * k.lumaT[ 1 ] + col[ 2 ] * k.lumaT[ 2 ] ) );
    for( int c = 0; c < 3; c++ )
    {
        out[ c ] = sat( lumaAdd * k.rcpLumaT[ c ] + col[ c ] );
    }
}
} // namespace

extern "C" void or_lpm_luma_coefficients( double out[ 3 ] )
{
    const float s[ 3 ] = { 0, 0, 0 }, c[ 3 ] = { 1, 1, 1 };
    const Lpm   k      = Setup( s, c );
    for( int i = 0; i < 3; i++ ) out[ i ] = k.lumaT[ i ];
}

extern "C" void or_lpm_map( const float saturation[ 3 ], const float crosstalk[ 3 ], const double in[ 3 ], double out[ 3 ] )
{
    Map( Setup( saturation, crosstalk ), in, out );
}

extern "C" void or_compose_final( const OrFramebuffers* f, const float saturation[ 3 ], const float crosstalk[ 3 ], double avgLuminance,
                                  double emissionMaxScreenColor )
{
    const Lpm k = Setup( saturation, crosstalk );
    const int W = f->width, H = f->height, sep = W / 2;
    const double* pre = f->cur[ RG_EXT_PLANE_PRE_FINAL ];
    const double* thr = f->cur[ RG_EXT_PLANE_THROUGHPUT ];
    double*       fin = f->cur[ RG_EXT_PLANE_FINAL ];
    const double* emi = f->cur[ RG_EXT_PLANE_SCREEN_EMISSION ];
    const double* fgp = f->cur[ RG_EXT_PLANE_ACID_FOG ];
    auto atp = [ & ]( const double* pl, int x, int y, int c ) { return ( x < 0 || y < 0 || x >= W || y >= H ) ? 0.0 : pl[ ( size_t( y ) * W + x ) * 4 + c ]; };
    auto at  = [ & ]( int x, int y, int c ) { return atp( pre, x, y, c ); };
    // ev100ToLuminousExposure(getAutoEV100()) (Exposure.h:29-56)
    const double lum      = std::max( 0.0, avgLuminance );
    const double maxLum   = 1.2 * std::exp2( std::log2( lum * 100.0 / 12.5 ) );
    const double exposure = maxLum > 0.0 ? 1.0 / maxLum : 0.0;
    for( int y = 0; y < H; y++ )
    {
        for( int x = 0; x < W; x++ )
        {
            int cx, cy;
            or_checkerboard_from_regular( x, y, W, &cx, &cy );
            const double* pl[ 3 ] = { pre, emi, fgp };
            double        v[ 3 ][ 3 ];
            for( int q = 0; q < 3; q++ )
                for( int c = 0; c < 3; c++ ) v[ q ][ c ] = atp( pl[ q ], x, y, c );
            if( thr[ ( size_t( cy ) * W + cx ) * 4 + 3 ] > 0.5 )
            {
                const double l = ( cx == 0 || cx == sep ) ? 0.0 : 1.0, r = ( cx == sep || cx == W ) ? 0.0 : 1.0;
                for( int q = 0; q < 3; q++ )
                {
                    for( int c = 0; c < 3; c++ )
                    {
                        const double cross = 0.25 * atp( pl[ q ], x - 1, y, c ) * l + 0.25 * atp( pl[ q ], x + 1, y, c ) * r +
                                             0.25 * atp( pl[ q ], x, y + 1, c ) + 0.25 * atp( pl[ q ], x, y - 1, c );
                        v[ q ][ c ]        = 0.5 * v[ q ][ c ] + 0.5 * cross;
                    }
                }
            }
            // finalizeColor (CmPrepareFinal.comp:109-121), then the acid fog (:162-164)
            double in[ 3 ], out[ 3 ];
            for( int c = 0; c < 3; c++ ) in[ c ] = v[ 0 ][ c ] * exposure + v[ 1 ][ c ] * emissionMaxScreenColor;
            Map( k, in, out );
            double* o = fin + ( size_t( y ) * W + x ) * 4;
            o[ 0 ] = out[ 0 ] + v[ 2 ][ 0 ], o[ 1 ] = out[ 1 ] + v[ 2 ][ 1 ], o[ 2 ] = out[ 2 ] + v[ 2 ][ 2 ], o[ 3 ] = 0.0;
        }
    }
}
