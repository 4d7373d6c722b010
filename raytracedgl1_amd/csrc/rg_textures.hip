#include "rg_textures.h"

#include <cmath>

namespace rg
{

namespace
{
constexpr int THREADS = 256;

// sRGB transfer functions (IEC 61966-2-1), as an sRGB format decodes on
// sampling and encodes on a blit.
double SrgbToLinear( double c )
{
    return c <= 0.04045 ? c / 12.92 : std::pow( ( c + 0.055 ) / 1.055, 2.4 );
}

__device__ uint32_t EncodeChannel( float c )
{
    c                = c < 0.0f ? 0.0f : ( c > 1.0f ? 1.0f : c );
    const float s    = c <= 0.0031308f ? c * 12.92f : 1.055f * powf( c, 1.0f / 2.4f ) - 0.055f;
    const float q    = rintf( s * 255.0f );
    return uint32_t( q < 0.0f ? 0.0f : ( q > 255.0f ? 255.0f : q ) );
}

// One level of the chain from the level above, as a linear-filtered blit of
// the whole image does: each destination texel centre maps into the source
// and is filtered bilinearly there (clamp to edge), in linear space.
__global__ __launch_bounds__( THREADS ) void k_mip_level( const uint32_t* src, int sw, int sh, uint32_t* dst, int dw, int dh, const float* lut,
                                                          uint32_t srgb )
{
    const int i = blockIdx.x * THREADS + threadIdx.x;
    if( i >= dw * dh )
    {
        return;
    }
    const int   dx = i % dw, dy = i / dw;
    const float x = ( float( dx ) + 0.5f ) * ( float( sw ) / float( dw ) ) - 0.5f;
    const float y = ( float( dy ) + 0.5f ) * ( float( sh ) / float( dh ) ) - 0.5f;
    const float fx = floorf( x ), fy = floorf( y );
    const float ax = x - fx, ay = y - fy;
    const int   x0 = int( fx ), y0 = int( fy );
    float       acc[ 4 ] = { 0.f, 0.f, 0.f, 0.f };
    for( int k = 0; k < 4; k++ )
    {
        const int   xx = x0 + ( k & 1 ), yy = y0 + ( k >> 1 );
        const int   cx = xx < 0 ? 0 : ( xx >= sw ? sw - 1 : xx ), cy = yy < 0 ? 0 : ( yy >= sh ? sh - 1 : yy );
        const float wgt = ( ( k & 1 ) ? ax : 1.0f - ax ) * ( ( k >> 1 ) ? ay : 1.0f - ay );
        const uint32_t c = src[ size_t( cy ) * sw + cx ];
        // UNORM channels (and alpha) filter the 8-bit codes themselves: with
        // the exact 1/4 weights of a halving blit the sums are exact, so the
        // rounding of .5 cases does not depend on the arithmetic's precision
        for( int ch = 0; ch < 3; ch++ )
        {
            const uint32_t b = ( c >> ( 8 * ch ) ) & 0xFFu;
            acc[ ch ] += wgt * ( srgb ? lut[ b ] : float( b ) );
        }
        acc[ 3 ] += wgt * float( c >> 24 );
    }
    auto code = []( float v ) { return uint32_t( rintf( v < 0.f ? 0.f : ( v > 255.f ? 255.f : v ) ) ); };
    uint32_t out = code( acc[ 3 ] ) << 24;
    for( int ch = 0; ch < 3; ch++ )
    {
        out |= ( srgb ? EncodeChannel( acc[ ch ] ) : code( acc[ ch ] ) ) << ( 8 * ch );
    }
    dst[ i ] = out;
}
} // namespace

void TextureManager::Provide( const RgOriginalTextureInfo& info, hipStream_t s )
{
    RG_REQUIRE( info.pTextureName && info.pTextureName[ 0 ], RG_RESULT_WRONG_FUNCTION_ARGUMENT, "texture name is empty" );
    RG_REQUIRE( info.pPixels, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "texture pixels are null" );
    RG_REQUIRE( info.size.width >= 1 && info.size.height >= 1 && info.size.width <= TEX_MAX_SIZE && info.size.height <= TEX_MAX_SIZE,
                RG_RESULT_WRONG_FUNCTION_ARGUMENT, "texture size must be in [1, 32768]" );
    if( !srgbLut.ptr )
    {
        float lut[ 256 ];
        for( int i = 0; i < 256; i++ )
        {
            lut[ i ] = float( SrgbToLinear( i / 255.0 ) );
        }
        srgbLut.Reserve( sizeof( lut ) );
        RG_HIP_CHECK( hipMemcpy( srgbLut.ptr, lut, sizeof( lut ), hipMemcpyHostToDevice ) );
    }
    const std::string name( info.pTextureName );
    auto endsWith = [ & ]( const char* sfx ) {
        const size_t n = std::strlen( sfx );
        return name.size() > n && name.compare( name.size() - n, n, sfx ) == 0;
    };
    // the occlusion-roughness-metallic and normal maps are UNORM, the others sRGB
    const uint32_t srgb = ( endsWith( "_orm" ) || endsWith( "_n" ) ) ? 0u : 1u;
    uint32_t          id;
    auto              it = byName.find( name );
    if( it != byName.end() )
    {
        id = it->second; // replaced in place
    }
    else if( !freeIds.empty() )
    {
        id = freeIds.back();
        freeIds.pop_back();
    }
    else
    {
        entries.push_back( std::make_unique< Entry >() );
        id = uint32_t( entries.size() );
    }
    Entry& e = *entries[ id - 1 ];

    const uint32_t w = info.size.width, h = info.size.height;
    uint32_t       mips = 1;
    size_t         total = size_t( w ) * h;
    for( uint32_t lw = w, lh = h; ( lw > 1 || lh > 1 ) && mips < TEX_MAX_MIPS; mips++ )
    {
        lw = lw > 1 ? lw >> 1 : 1;
        lh = lh > 1 ? lh >> 1 : 1;
        total += size_t( lw ) * lh;
    }
    e.texels.Release();
    e.texels.Reserve( total * 4 );
    RG_HIP_CHECK( hipMemcpyAsync( e.texels.ptr, info.pPixels, size_t( w ) * h * 4, hipMemcpyHostToDevice, s ) );
    uint32_t* level = e.texels.as< uint32_t >();
    for( uint32_t l = 1, lw = w, lh = h; l < mips; l++ )
    {
        const uint32_t nw = lw > 1 ? lw >> 1 : 1, nh = lh > 1 ? lh >> 1 : 1;
        uint32_t*      next = level + size_t( lw ) * lh;
        k_mip_level<<< DivUp( nw * nh, THREADS ), THREADS, 0, s >>>( level, int( lw ), int( lh ), next, int( nw ), int( nh ), srgbLut.as< float >(),
                                                                     srgb );
        CheckLaunch( "k_mip_level" );
        level = next, lw = nw, lh = nh;
    }
    RG_HIP_CHECK( hipStreamSynchronize( s ) ); // the caller's pixels may go away on return

    // the reference's filter choice (SamplerManager.cpp:30-40, :245-250):
    // AUTO is the default dynamic filter, linear
    const bool nearest = info.filter == RG_SAMPLER_FILTER_NEAREST;
    e.desc.texels      = e.texels.as< uint32_t >();
    e.desc.width       = w;
    e.desc.height      = h;
    e.desc.mips        = mips;
    e.desc.magNearest  = nearest ? 1u : 0u;
    e.desc.minNearest  = ( nearest && !forceMinLinear ) ? 1u : 0u;
    e.desc.clampU      = info.addressModeU == RG_SAMPLER_ADDRESS_MODE_CLAMP ? 1u : 0u;
    e.desc.clampV      = info.addressModeV == RG_SAMPLER_ADDRESS_MODE_CLAMP ? 1u : 0u;
    e.desc.srgb        = srgb;
    e.live             = true;
    byName[ name ]     = id;
    dirty              = true;
}

void TextureManager::MarkDeleted( const char* name )
{
    RG_REQUIRE( name, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "texture name is null" );
    auto it = byName.find( name );
    if( it == byName.end() )
    {
        return;
    }
    Entry& e = *entries[ it->second - 1 ];
    e.texels.Release();
    e.desc = TextureDesc{};
    e.live = false;
    freeIds.push_back( it->second );
    byName.erase( it );
    dirty = true;
}

uint32_t TextureManager::Find( const char* name ) const
{
    if( !name || !name[ 0 ] )
    {
        return TEX_NONE;
    }
    auto it = byName.find( name );
    return it == byName.end() ? TEX_NONE : it->second;
}

TexView TextureManager::View( hipStream_t s )
{
    if( dirty && !entries.empty() )
    {
        hostTable.resize( entries.size() );
        for( size_t i = 0; i < entries.size(); i++ )
        {
            hostTable[ i ] = entries[ i ]->desc;
        }
        table.Reserve( hostTable.size() * sizeof( TextureDesc ) );
        RG_HIP_CHECK( hipMemcpyAsync( table.ptr, hostTable.data(), hostTable.size() * sizeof( TextureDesc ), hipMemcpyHostToDevice, s ) );
        RG_HIP_CHECK( hipStreamSynchronize( s ) );
        dirty = false;
    }
    TexView v;
    v.descs = table.as< TextureDesc >();
    v.count = uint32_t( entries.size() );
    v.srgb  = srgbLut.as< float >();
    return v;
}

} // namespace rg
