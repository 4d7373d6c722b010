// Luminance histogram and adaptive average luminance (eye adaptation).
//
// CmLuminanceHistogram.comp:41-86 builds a 256-bin histogram per 16x16
// group in shared memory and adds it to the global one; here a persistent
// grid of 256-thread blocks walks the image with a grid stride, so each block
// folds many pixels into its LDS bins before one global atomic per bin, and
// the lanes of a wave that share a bin add to it with one LDS atomic.
// CmLuminanceAvg.comp:39-90 is a single 256-thread group.
#include "rg_device_math.h"
#include "rg_framebuffers.h"
#include "rg_internal.h"
#include "rg_kernels.h"

#include <cmath>

namespace rg
{
namespace
{
constexpr int HIST_THREADS = 256;
constexpr int HIST_BINS    = 256;

// getColorBin (CmLuminanceHistogram.comp:41-54), with a deterministic log2.
// NaN luminance goes to bin 0, +inf to bin 255.
__device__ __forceinline__ uint32_t ColorBin( float lum, float minLog, float invRange )
{
    if( !( lum >= 0.001f ) )
    {
        return 0;
    }
    if( lum > 3.0e38f )
    {
        return 255;
    }
    const float l = clampf( ( rgLog2( lum ) - minLog ) * invRange, 0.0f, 1.0f );
    return uint32_t( l * 254.0f + 1.0f );
}

__global__ __launch_bounds__( HIST_THREADS ) void k_lum_histogram( const float* lum, uint32_t count, float minLog, float maxLog, uint32_t* hist )
{
    __shared__ uint32_t s_bins[ HIST_BINS ];
    s_bins[ threadIdx.x ] = 0;
    __syncthreads();
    const float    invRange = 1.0f / ( maxLog - minLog );
    const uint32_t lane     = threadIdx.x & 63u;
    for( uint32_t i = blockIdx.x * HIST_THREADS + threadIdx.x; i < count; i += gridDim.x * HIST_THREADS )
    {
        // Wave-ballot aggregation: the lanes of a wave holding the same bin
        // add their count with one LDS atomic from their first lane (image
        // regions of one luminance put a whole wave into a few bins).
        const uint32_t bin     = ColorBin( lum[ i ], minLog, invRange );
        bool           pending = true;
        for( ;; )
        {
            const uint64_t act = __ballot( pending );
            if( act == 0ull )
            {
                break;
            }
            const int      leader = __ffsll( ( long long )act ) - 1;
            const uint32_t b      = __shfl( bin, leader, 64 );
            const uint64_t same   = __ballot( pending && bin == b );
            if( int( lane ) == leader )
            {
                atomicAdd( &s_bins[ b ], uint32_t( __popcll( same ) ) );
            }
            pending = pending && bin != b;
        }
    }
    __syncthreads();
    const uint32_t c = s_bins[ threadIdx.x ];
    if( c != 0 )
    {
        atomicAdd( &hist[ threadIdx.x ], c );
    }
}

__global__ __launch_bounds__( HIST_BINS ) void k_lum_average( uint32_t* hist,
                                                              uint32_t  pixelCount,
                                                              float     minLog,
                                                              float     maxLog,
                                                              float     timeDelta,
                                                              uint32_t  stop,
                                                              float*    avgLum )
{
    __shared__ uint32_t s[ HIST_BINS ];
    const uint32_t      i     = threadIdx.x;
    const uint32_t      count = hist[ i ];
    s[ i ]                    = count * i;
    __syncthreads();
    hist[ i ] = 0; // cleared for the next frame
    for( uint32_t cutoff = HIST_BINS >> 1; cutoff > 0; cutoff >>= 1 )
    {
        if( i < cutoff )
        {
            s[ i ] += s[ i + cutoff ];
        }
        __syncthreads();
    }
    if( i == 0 && stop == 0 )
    {
        const float range     = maxLog - minLog;
        const float pixels    = float( pixelCount );
        const float black     = float( count );
        const float weighted  = ( float( s[ 0 ] ) / fmaxf( pixels - black, 1.0f ) ) - 1.0f;
        const float avg       = rgExp2( weighted / 254.0f * range + minLog );
        const float last      = *avgLum;
        // t = 1 - exp(-dt * 1.1), exp(x) = 2^(x log2 e)
        const float t         = 1.0f - rgExp2( -timeDelta * 1.1f * 1.44269504f );
        const float adapted   = last + ( avg - last ) * t;
        *avgLum               = fminf( adapted, rgExp2( maxLog ) );
    }
}

int HistogramGrid()
{
    static int blocks = 0;
    if( blocks == 0 )
    {
        int dev = 0, cus = 0;
        ( void )hipGetDevice( &dev );
        ( void )hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev );
        blocks = ( cus > 0 ? cus : 64 ) * 4;
    }
    return blocks;
}
// ------------------------------------------------------------ final compose

constexpr int CBX = 32, CBY = 8;

// LPM 709 -> 709 filter (LpmMap, Shaders/LPM/ffx_lpm.h:726-828, with
// shoulder, con, soft and con2 off: LPM_CONFIG_709_709).
__device__ __forceinline__ v3 LpmMap709( v3 c, const LpmConsts& k )
{
    const float mx = fmaxf( c.x, fmaxf( c.y, c.z ) );
    if( !( mx > 0.0f ) )
    {
        return mk3( 0, 0, 0 );
    }
    const float rcpMax = 1.0f / mx;
    v3          ratio  = mk3( powf( c.x * rcpMax, k.saturation[ 0 ] ), powf( c.y * rcpMax, k.saturation[ 1 ] ), powf( c.z * rcpMax, k.saturation[ 2 ] ) );
    float       luma   = c.y * k.lumaT[ 1 ] + ( c.x * k.lumaT[ 0 ] + ( c.z * k.lumaT[ 2 ] ) );
    luma               = powf( luma, k.contrast );
    luma               = luma / ( luma * k.toneScaleBias[ 0 ] + k.toneScaleBias[ 1 ] );
    const float lumaRatio  = ratio.x * k.lumaT[ 0 ] + ratio.y * k.lumaT[ 1 ] + ratio.z * k.lumaT[ 2 ];
    const float ratioScale = saturate( luma / lumaRatio );
    v3          col        = mk3( saturate( ratio.x * ratioScale ), saturate( ratio.y * ratioScale ), saturate( ratio.z * ratioScale ) );
    // crosstalk: push the unclipped channels up to make up for clipping
    const v3    cap     = mk3( k.crosstalk[ 0 ] - k.crosstalk[ 0 ] * col.x, k.crosstalk[ 1 ] - k.crosstalk[ 1 ] * col.y, k.crosstalk[ 2 ] - k.crosstalk[ 2 ] * col.z );
    float       lumaAdd = saturate( luma - ( col.y * k.lumaT[ 1 ] + ( col.x * k.lumaT[ 0 ] + col.z * k.lumaT[ 2 ] ) ) );
    const float t       = lumaAdd / ( cap.y * k.lumaT[ 1 ] + ( cap.x * k.lumaT[ 0 ] + cap.z * k.lumaT[ 2 ] ) );
    col                 = mk3( saturate( t * cap.x + col.x ), saturate( t * cap.y + col.y ), saturate( t * cap.z + col.z ) );
    lumaAdd             = saturate( luma - ( col.z * k.lumaT[ 2 ] + ( col.x * k.lumaT[ 0 ] + col.y * k.lumaT[ 1 ] ) ) );
    return mk3( saturate( lumaAdd * k.rcpLumaT[ 0 ] + col.x ), saturate( lumaAdd * k.rcpLumaT[ 1 ] + col.y ), saturate( lumaAdd * k.rcpLumaT[ 2 ] + col.z ) );
}

__device__ __forceinline__ float4 LdReg( const float4* p, int W, int H, int x, int y )
{
    return ( x < 0 || y < 0 || x >= W || y >= H ) ? make_float4( 0.f, 0.f, 0.f, 0.f ) : p[ size_t( y ) * W + x ];
}

// CmCheckerboard.comp:49-76 then CmPrepareFinal.comp:148-176 (no volumetrics;
// classic shading is off at the default lightmapScreenCoverage of 0)
__global__ __launch_bounds__( CBX * CBY ) void k_compose( FbPtrs P, int W, int H, LpmConsts k, float emisWeight, const float* avgLum )
{
    const int x = blockIdx.x * CBX + threadIdx.x % CBX, y = blockIdx.y * CBY + threadIdx.x / CBX;
    if( x >= W || y >= H )
    {
        return;
    }
    const float4* pre = reinterpret_cast< const float4* >( P.cur[ RG_EXT_PLANE_PRE_FINAL ] );
    float4        c4  = pre[ size_t( y ) * W + x ];
    int           cx, cy;
    cbFromRegular( x, y, W, &cx, &cy );
    const float4 thr = reinterpret_cast< const float4* >( P.cur[ RG_EXT_PLANE_THROUGHPUT ] )[ size_t( cy ) * W + cx ];
    const float4* em  = reinterpret_cast< const float4* >( P.cur[ RG_EXT_PLANE_SCREEN_EMISSION ] );
    const float4* fg  = reinterpret_cast< const float4* >( P.cur[ RG_EXT_PLANE_ACID_FOG ] );
    const float4  e4  = em[ size_t( y ) * W + x ], f4 = fg[ size_t( y ) * W + x ];
    v3           hdr = mk3( c4.x, c4.y, c4.z ), emis = mk3( e4.x, e4.y, e4.z ), fog = mk3( f4.x, f4.y, f4.z );
    if( thr.w > 0.5f ) // wasSplit: a refracted / reflected pixel is resolved with its cross
    {
        const int   sep = cbSeparator( W );
        const float l   = ( cx == 0 || cx == sep ) ? 0.0f : 1.0f;
        const float r   = ( cx == sep || cx == W ) ? 0.0f : 1.0f;
        auto resolve = [ & ]( const float4* src, v3 center ) {
            const float4 a = LdReg( src, W, H, x - 1, y ), b = LdReg( src, W, H, x + 1, y ), u = LdReg( src, W, H, x, y + 1 ), d = LdReg( src, W, H, x, y - 1 );
            const v3 cross = mk3( a.x, a.y, a.z ) * ( 0.25f * l ) + mk3( b.x, b.y, b.z ) * ( 0.25f * r ) + mk3( u.x, u.y, u.z ) * 0.25f + mk3( d.x, d.y, d.z ) * 0.25f;
            return mix3( center, cross, 0.5f );
        };
        hdr  = resolve( pre, hdr );
        emis = resolve( em, emis );
        fog  = resolve( fg, fog );
    }
    // ev100ToLuminousExposure(getAutoEV100()) (Exposure.h:29-56)
    const float lum      = fmaxf( 0.0f, *avgLum );
    const float ev100    = log2f( lum * 100.0f / 12.5f );
    const float maxLum   = 1.2f * exp2f( ev100 );
    const float exposure = maxLum > 0.0f ? 1.0f / maxLum : 0.0f;
    // finalizeColor: hdr * exposure + screen emission * emissionMaxScreenColor,
    // tone mapped; then the acid fog (CmPrepareFinal.comp:109-121, :162-164)
    const v3    out      = LpmMap709( hdr * exposure + emis * emisWeight, k ) + fog;
    reinterpret_cast< float4* >( P.cur[ RG_EXT_PLANE_FINAL ] )[ size_t( y ) * W + x ] = make_float4( out.x, out.y, out.z, 0.0f );
}

} // namespace

// LpmSetup (Shaders/LPM/LpmSetupCustom.inl, ffx_lpm.h:300-420) for the
// reference's call (ImageComposition.cpp:377-394): LPM_CONFIG_709_709,
// LPM_COLORS_709_709, softGap 0, hdrMax 256, exposure 8, contrast 0.1,
// shoulder contrast 1; computed in FP64 and stored as FP32.
LpmConsts MakeLpmConsts( const float saturation[ 3 ], const float crosstalk[ 3 ] )
{
    const double hdrMax = 256.0, exposure = 8.0, shoulderContrast = 1.0;
    const double contrast = 0.1 + 1.0;
    const double midIn = hdrMax * 0.18 * std::exp2( -exposure ), midOut = 0.18;
    const double cs = contrast * shoulderContrast;
    const double z0 = -std::pow( midIn, contrast );
    const double z1 = std::pow( hdrMax, cs ) * std::pow( midIn, contrast );
    const double z2 = std::pow( hdrMax, contrast ) * std::pow( midIn, cs ) * midOut;
    const double z3 = std::pow( hdrMax, cs ) * midOut;
    const double z4 = std::pow( midIn, cs ) * midOut;
    const double tsb0 = -( ( z0 + ( midOut * ( z1 - z2 ) ) / ( z3 - z4 ) ) / z4 );
    const double w0 = std::pow( hdrMax, cs ) * std::pow( midIn, contrast );
    const double w1 = std::pow( hdrMax, contrast ) * std::pow( midIn, cs ) * midOut;
    const double w2 = std::pow( hdrMax, cs ) * midOut;
    const double w3 = std::pow( midIn, cs ) * midOut;
    const double tsb1 = ( w0 - w1 ) / ( w2 - w3 );
    // Y row of the Rec.709 / D65 RGB -> XYZ matrix (LpmColRgbToXyz), normalised
    const double xy[ 4 ][ 2 ] = { { 0.64, 0.33 }, { 0.30, 0.60 }, { 0.15, 0.06 }, { 0.3127, 0.3290 } };
    double       m[ 3 ][ 3 ]; // columns: primaries' xyz
    for( int c = 0; c < 3; c++ )
    {
        m[ 0 ][ c ] = xy[ c ][ 0 ];
        m[ 1 ][ c ] = xy[ c ][ 1 ];
        m[ 2 ][ c ] = 1.0 - xy[ c ][ 0 ] - xy[ c ][ 1 ];
    }
    const double wX = xy[ 3 ][ 0 ] / xy[ 3 ][ 1 ], wY = 1.0, wZ = ( 1.0 - xy[ 3 ][ 0 ] - xy[ 3 ][ 1 ] ) / xy[ 3 ][ 1 ];
    const double det = m[ 0 ][ 0 ] * ( m[ 1 ][ 1 ] * m[ 2 ][ 2 ] - m[ 1 ][ 2 ] * m[ 2 ][ 1 ] ) - m[ 0 ][ 1 ] * ( m[ 1 ][ 0 ] * m[ 2 ][ 2 ] - m[ 1 ][ 2 ] * m[ 2 ][ 0 ] ) +
                       m[ 0 ][ 2 ] * ( m[ 1 ][ 0 ] * m[ 2 ][ 1 ] - m[ 1 ][ 1 ] * m[ 2 ][ 0 ] );
    double inv[ 3 ][ 3 ];
    inv[ 0 ][ 0 ] = ( m[ 1 ][ 1 ] * m[ 2 ][ 2 ] - m[ 1 ][ 2 ] * m[ 2 ][ 1 ] ) / det;
    inv[ 0 ][ 1 ] = ( m[ 0 ][ 2 ] * m[ 2 ][ 1 ] - m[ 0 ][ 1 ] * m[ 2 ][ 2 ] ) / det;
    inv[ 0 ][ 2 ] = ( m[ 0 ][ 1 ] * m[ 1 ][ 2 ] - m[ 0 ][ 2 ] * m[ 1 ][ 1 ] ) / det;
    inv[ 1 ][ 0 ] = ( m[ 1 ][ 2 ] * m[ 2 ][ 0 ] - m[ 1 ][ 0 ] * m[ 2 ][ 2 ] ) / det;
    inv[ 1 ][ 1 ] = ( m[ 0 ][ 0 ] * m[ 2 ][ 2 ] - m[ 0 ][ 2 ] * m[ 2 ][ 0 ] ) / det;
    inv[ 1 ][ 2 ] = ( m[ 0 ][ 2 ] * m[ 1 ][ 0 ] - m[ 0 ][ 0 ] * m[ 1 ][ 2 ] ) / det;
    inv[ 2 ][ 0 ] = ( m[ 1 ][ 0 ] * m[ 2 ][ 1 ] - m[ 1 ][ 1 ] * m[ 2 ][ 0 ] ) / det;
    inv[ 2 ][ 1 ] = ( m[ 0 ][ 1 ] * m[ 2 ][ 0 ] - m[ 0 ][ 0 ] * m[ 2 ][ 1 ] ) / det;
    inv[ 2 ][ 2 ] = ( m[ 0 ][ 0 ] * m[ 1 ][ 1 ] - m[ 0 ][ 1 ] * m[ 1 ][ 0 ] ) / det;
    double scale[ 3 ], lumaT[ 3 ];
    for( int c = 0; c < 3; c++ )
    {
        scale[ c ] = inv[ c ][ 0 ] * wX + inv[ c ][ 1 ] * wY + inv[ c ][ 2 ] * wZ;
        lumaT[ c ] = m[ 1 ][ c ] * scale[ c ];
    }
    const double sum = lumaT[ 0 ] + lumaT[ 1 ] + lumaT[ 2 ];
    LpmConsts    k;
    for( int c = 0; c < 3; c++ )
    {
        lumaT[ c ] /= sum;
        k.lumaT[ c ]      = float( lumaT[ c ] );
        k.rcpLumaT[ c ]   = float( 1.0 / lumaT[ c ] );
        k.saturation[ c ] = float( double( saturation[ c ] ) + contrast );
        k.crosstalk[ c ]  = crosstalk[ c ];
    }
    k.contrast          = float( contrast );
    k.toneScaleBias[ 0 ] = float( tsb0 );
    k.toneScaleBias[ 1 ] = float( tsb1 );
    return k;
}

void ComposeFinal( FramebufferSet& fb, const LpmConsts& k, float emissionMaxScreenColor, const float* d_avgLum, hipStream_t stream )
{
    const dim3 grid( DivUp( fb.width, CBX ), DivUp( fb.height, CBY ) );
    k_compose<<< grid, CBX * CBY, 0, stream >>>( fb.Ptrs(), int( fb.width ), int( fb.height ), k, emissionMaxScreenColor, d_avgLum );
    CheckLaunch( "k_compose" );
}

namespace
{
// FINAL (float RGBA, regular layout) -> RGBA8 with the sRGB transfer of the
// swapchain's *_SRGB formats (IEC 61966-2-1), R in the low byte, alpha 255.
__global__ __launch_bounds__( 256 ) void k_pack_rgba8( const float4* src, uint32_t* dst, uint32_t n )
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if( i >= n )
    {
        return;
    }
    const float4 c   = src[ i ];
    auto         enc = []( float v ) {
        v             = fminf( fmaxf( v, 0.0f ), 1.0f );
        const float e = v <= 0.0031308f ? 12.92f * v : 1.055f * powf( v, 1.0f / 2.4f ) - 0.055f;
        return uint32_t( fminf( fmaxf( e, 0.0f ), 1.0f ) * 255.0f + 0.5f );
    };
    dst[ i ] = enc( c.x ) | ( enc( c.y ) << 8 ) | ( enc( c.z ) << 16 ) | ( 255u << 24 );
}
} // namespace

void PackFinalRgba8( const float* d_final, uint32_t* d_out, uint32_t pixels, hipStream_t stream )
{
    if( pixels == 0 )
    {
        return;
    }
    k_pack_rgba8<<< DivUp( pixels, 256u ), 256, 0, stream >>>( reinterpret_cast< const float4* >( d_final ), d_out, pixels );
    CheckLaunch( "k_pack_rgba8" );
}

void LuminanceHistogram( const float* d_lum, uint32_t count, float minLog, float maxLog, uint32_t* d_hist, hipStream_t stream )
{
    RG_REQUIRE( maxLog > minLog, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "maxLogLuminance must exceed minLogLuminance" );
    if( count == 0 )
    {
        return;
    }
    const uint32_t blocks = DivUp( count, HIST_THREADS ) < uint32_t( HistogramGrid() ) ? DivUp( count, HIST_THREADS ) : uint32_t( HistogramGrid() );
    k_lum_histogram<<< blocks, HIST_THREADS, 0, stream >>>( d_lum, count, minLog, maxLog, d_hist );
    CheckLaunch( "k_lum_histogram" );
}

void LuminanceAverage( uint32_t* d_hist, uint32_t pixelCount, float minLog, float maxLog, float timeDelta, bool stop, float* d_avg, hipStream_t stream )
{
    k_lum_average<<< 1, HIST_BINS, 0, stream >>>( d_hist, pixelCount, minLog, maxLog, timeDelta, stop ? 1u : 0u, d_avg );
    CheckLaunch( "k_lum_average" );
}

} // namespace rg
