// Albedo textures of rgProvideOriginalTexture (Source/TextureManager.cpp:
// 398-440): R8G8B8A8 sRGB texels with a full mip chain made on the device
// (the linear blits of TextureUploader.cpp:104-175), sampled in the
// traversal's hit shading with the reference's sampler set
// (SamplerManager.cpp:213-265: nearest or linear per texture, linear between
// mips, repeat or clamp-to-edge per axis).  Filtering is isotropic.
#pragma once

#include "rg_internal.h"
#include "rg_device_math.h"

#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace rg
{

// Texture ids in geometry records are table index + 1; 0 is no texture.
// A material named N takes its maps from the textures named N (albedo, sRGB),
// N_orm (occlusion-roughness-metallic, UNORM), N_n (normal, UNORM) and N_e
// (emission, sRGB): the reference's override postfixes (Source/Const.h:40-43,
// TextureManager.cpp:418-425), applied to provided textures.
constexpr uint32_t TEX_NONE      = 0;
constexpr uint32_t TEX_MAX_MIPS  = 16;
constexpr uint32_t TEX_MAX_SIZE  = 1u << 15;

struct TextureDesc
{
    const uint32_t* texels; // RGBA8 (sRGB colour, linear alpha), mip 0 first, levels back to back
    uint32_t        width, height, mips;
    uint32_t        magNearest, minNearest; // 1: nearest texel within a level
    uint32_t        clampU, clampV;         // 1: clamp to edge, 0: repeat
    uint32_t        srgb;                   // 0: UNORM colour (the _orm and _n maps)
};

// What the kernels see of the texture table.
struct TexView
{
    const TextureDesc* descs  = nullptr;
    uint32_t           count  = 0;
    const float*       srgb   = nullptr; // 256-entry sRGB -> linear table
};

// Per-triangle record of an alpha-tested, textured primitive
// (RtAlphaTest.rahit): texture id, layer-0 colour and the three texcoords.
struct AlphaRec
{
    uint32_t tex;
    uint32_t color;
    float    uv[ 6 ];
};
static_assert( sizeof( AlphaRec ) == 32, "AlphaRec" );

struct TextureManager
{
    struct Entry
    {
        DeviceBuffer texels;
        TextureDesc  desc{};
        bool         live = false;
    };
    std::unordered_map< std::string, uint32_t > byName; // name -> id
    std::vector< std::unique_ptr< Entry > >     entries;  // id - 1
    std::vector< uint32_t >                     freeIds;
    DeviceBuffer                                table, srgbLut;
    std::vector< TextureDesc >                  hostTable;
    bool                                        dirty = true;
    bool                                        forceMinLinear = false;

    // Creates or replaces `name`; builds the mip chain on `s` and waits for it.
    void     Provide( const RgOriginalTextureInfo& info, hipStream_t s );
    void     MarkDeleted( const char* name );
    uint32_t Find( const char* name ) const;
    // Uploads the table if it changed; valid until the next Provide / MarkDeleted.
    TexView  View( hipStream_t s );
};

// ------------------------------------------------------------ sampling

// A texture id that names a live texture of the table (a deleted texture's
// slot stays in the table with no texels).
__device__ __forceinline__ bool TexValid( const TexView& tv, uint32_t id )
{
    return id != TEX_NONE && id <= tv.count && tv.descs[ id - 1 ].texels != nullptr;
}

__device__ __forceinline__ v3 rgbOf( float4 c )
{
    return mk3( c.x, c.y, c.z );
}

__device__ __forceinline__ int TexWrap( int i, int n, uint32_t clampMode )
{
    if( clampMode )
    {
        return i < 0 ? 0 : ( i >= n ? n - 1 : i );
    }
    const int m = i % n;
    return m < 0 ? m + n : m;
}

__device__ __forceinline__ float4 TexelAt( const TexView& tv, const TextureDesc& d, const uint32_t* level, int w, int h, int x, int y )
{
    const uint32_t c = level[ size_t( TexWrap( y, h, d.clampV ) ) * w + TexWrap( x, w, d.clampU ) ];
    if( !d.srgb )
    {
        return make_float4( float( c & 0xFFu ) * ( 1.0f / 255.0f ), float( ( c >> 8 ) & 0xFFu ) * ( 1.0f / 255.0f ),
                            float( ( c >> 16 ) & 0xFFu ) * ( 1.0f / 255.0f ), float( c >> 24 ) * ( 1.0f / 255.0f ) );
    }
    return make_float4( tv.srgb[ c & 0xFFu ], tv.srgb[ ( c >> 8 ) & 0xFFu ], tv.srgb[ ( c >> 16 ) & 0xFFu ], float( c >> 24 ) * ( 1.0f / 255.0f ) );
}

// One mip level, nearest or bilinear (texel centres at i + 0.5).
__device__ float4 SampleLevel( const TexView& tv, const TextureDesc& d, uint32_t lvl, float u, float v, bool nearest )
{
    const uint32_t* base = d.texels;
    int             w = int( d.width ), h = int( d.height );
    for( uint32_t k = 0; k < lvl; k++ )
    {
        base += size_t( w ) * h;
        w = w > 1 ? w >> 1 : 1;
        h = h > 1 ? h >> 1 : 1;
    }
    if( nearest )
    {
        return TexelAt( tv, d, base, w, h, int( floorf( u * float( w ) ) ), int( floorf( v * float( h ) ) ) );
    }
    const float x = u * float( w ) - 0.5f, y = v * float( h ) - 0.5f;
    const float fx = floorf( x ), fy = floorf( y );
    const float ax = x - fx, ay = y - fy;
    const int   x0 = int( fx ), y0 = int( fy );
    const float4 t00 = TexelAt( tv, d, base, w, h, x0, y0 ), t10 = TexelAt( tv, d, base, w, h, x0 + 1, y0 );
    const float4 t01 = TexelAt( tv, d, base, w, h, x0, y0 + 1 ), t11 = TexelAt( tv, d, base, w, h, x0 + 1, y0 + 1 );
    auto lerp4 = []( float4 a, float4 b, float t ) {
        return make_float4( a.x + ( b.x - a.x ) * t, a.y + ( b.y - a.y ) * t, a.z + ( b.z - a.z ) * t, a.w + ( b.w - a.w ) * t );
    };
    return lerp4( lerp4( t00, t10, ax ), lerp4( t01, t11, ax ), ay );
}

// textureLod: magnification at lod <= 0, else linear between the two nearest
// levels (VK_SAMPLER_MIPMAP_MODE_LINEAR), lod clamped to the chain.
__device__ float4 SampleLod( const TexView& tv, uint32_t id, float u, float v, float lod )
{
    const TextureDesc& d = tv.descs[ id - 1 ];
    if( !( lod > 0.0f ) )
    {
        return SampleLevel( tv, d, 0, u, v, d.magNearest != 0 );
    }
    const float    q  = float( d.mips - 1 );
    const float    l  = lod < q ? lod : q;
    const uint32_t l0 = uint32_t( floorf( l ) );
    const uint32_t l1 = l0 + 1 < d.mips ? l0 + 1 : l0;
    const float    t  = l - float( l0 );
    const float4   a  = SampleLevel( tv, d, l0, u, v, d.minNearest != 0 );
    if( t == 0.0f || l1 == l0 )
    {
        return a;
    }
    const float4 b = SampleLevel( tv, d, l1, u, v, d.minNearest != 0 );
    return make_float4( a.x + ( b.x - a.x ) * t, a.y + ( b.y - a.y ) * t, a.z + ( b.z - a.z ) * t, a.w + ( b.w - a.w ) * t );
}

// textureGrad: lod = log2 of the longer of the two scaled gradient vectors.
__device__ float4 SampleGrad( const TexView& tv, uint32_t id, float u, float v, float dudx, float dvdx, float dudy, float dvdy )
{
    const TextureDesc& d  = tv.descs[ id - 1 ];
    const float        W  = float( d.width ), H = float( d.height );
    const float        rx = sqrtf( dudx * W * dudx * W + dvdx * H * dvdx * H );
    const float        ry = sqrtf( dudy * W * dudy * W + dvdy * H * dvdy * H );
    const float        rho = rx > ry ? rx : ry;
    return SampleLod( tv, id, u, v, rho > 0.0f ? log2f( rho ) : -INFINITY );
}

} // namespace rg
