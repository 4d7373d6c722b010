#include "rg_framebuffers.h"

namespace rg
{

namespace
{
// Same order as RgExtPlane.
const PlaneDesc g_planes[ RG_EXT_PLANE_COUNT ] = {
    { LAYOUT_CHECKERBOARD, 1, true },  // NORMAL
    { LAYOUT_CHECKERBOARD, 2, true },  // METALLIC_ROUGHNESS
    { LAYOUT_CHECKERBOARD, 1, true },  // DEPTH_WORLD
    { LAYOUT_CHECKERBOARD, 1, false }, // DEPTH_GRAD
    { LAYOUT_CHECKERBOARD, 4, false }, // MOTION
    { LAYOUT_CHECKERBOARD, 4, false }, // UNFILTERED_DIRECT
    { LAYOUT_CHECKERBOARD, 4, false }, // UNFILTERED_SPECULAR
    { LAYOUT_CHECKERBOARD, 4, false }, // UNFILTERED_INDIR
    { LAYOUT_CHECKERBOARD, 4, true },  // SURFACE_POSITION
    { LAYOUT_CHECKERBOARD, 4, true },  // VISIBILITY_BUFFER
    { LAYOUT_CHECKERBOARD, 4, false }, // VIEW_DIRECTION
    { LAYOUT_CHECKERBOARD, 4, false }, // THROUGHPUT
    { LAYOUT_CHECKERBOARD, 1, false }, // IS_SKY
    { LAYOUT_CHECKERBOARD, 4, true },  // ACCUM_HISTORY_LENGTH
    { LAYOUT_CHECKERBOARD, 4, false }, // DIFF_TEMPORARY
    { LAYOUT_CHECKERBOARD, 4, false }, // DIFF_ACCUM_COLOR
    { LAYOUT_CHECKERBOARD, 2, true },  // DIFF_ACCUM_MOMENTS
    { LAYOUT_CHECKERBOARD, 4, false }, // DIFF_COLOR_HISTORY
    { LAYOUT_CHECKERBOARD, 4, false }, // DIFF_PING
    { LAYOUT_CHECKERBOARD, 4, false }, // DIFF_PONG
    { LAYOUT_CHECKERBOARD, 4, true },  // SPEC_ACCUM_COLOR
    { LAYOUT_CHECKERBOARD, 4, false }, // SPEC_PING
    { LAYOUT_CHECKERBOARD, 4, false }, // SPEC_PONG
    { LAYOUT_CHECKERBOARD, 4, true },  // INDIR_ACCUM
    { LAYOUT_CHECKERBOARD, 4, false }, // INDIR_PING
    { LAYOUT_CHECKERBOARD, 4, false }, // INDIR_PONG
    { LAYOUT_CHECKERBOARD, 1, false }, // ATROUS_FILTERED_VARIANCE
    { LAYOUT_CHECKERBOARD, 4, true },  // RESERVOIRS
    { LAYOUT_CHECKERBOARD, 4, false }, // RESERVOIRS_INITIAL
    { LAYOUT_CHECKERBOARD, 2, true },  // GRADIENT_INPUTS
    { LAYOUT_REGULAR, 4, false },      // ALBEDO
    { LAYOUT_REGULAR, 1, false },      // HISTOGRAM_INPUT
    { LAYOUT_REGULAR, 4, false },      // PRE_FINAL
    { LAYOUT_STRATA, 4, false },       // DIS_PING_GRADIENT
    { LAYOUT_STRATA, 4, false },       // DIS_PONG_GRADIENT
    { LAYOUT_STRATA, 4, false },       // DIS_GRADIENT_HISTORY
    { LAYOUT_STRATA, 1, false },       // GRADIENT_PREV_PIX
    { LAYOUT_CHECKERBOARD, 4, false }, // INDIR_INITIAL_SAMPLE
    { LAYOUT_CHECKERBOARD, 2, false }, // INDIR_INITIAL_RADIANCE
    { LAYOUT_CHECKERBOARD, 4, true },  // INDIR_RESERVOIR_SAMPLE
    { LAYOUT_CHECKERBOARD, 4, true },  // INDIR_RESERVOIR_WEIGHT
    { LAYOUT_REGULAR, 4, false },      // FINAL
    { LAYOUT_REGULAR, 4, false },      // SCREEN_EMISSION
    { LAYOUT_REGULAR, 4, false },      // ACID_FOG
    { LAYOUT_CHECKERBOARD, 4, false }, // DENOISE_GUIDE
    { LAYOUT_CHECKERBOARD, 16, false }, // GI_REUSE
};
} // namespace

const PlaneDesc& GetPlaneDesc( int plane )
{
    return g_planes[ plane ];
}

bool IsPersistentPlane( int plane )
{
    return g_planes[ plane ].prev || plane == RG_EXT_PLANE_DIFF_COLOR_HISTORY || plane == RG_EXT_PLANE_DIS_GRADIENT_HISTORY ||
           plane == RG_EXT_PLANE_GRADIENT_PREV_PIX || plane == RG_EXT_PLANE_FINAL;
}

namespace
{
size_t PlaneBytesFor( int plane, uint32_t w, uint32_t h, uint32_t storage )
{
    const PlaneDesc& d  = g_planes[ plane ];
    const size_t     gw = ( w + 1 ) / 3 > 1 ? ( w + 1 ) / 3 : 1, gh = ( h + 1 ) / 3 > 1 ? ( h + 1 ) / 3 : 1;
    const size_t     px = d.layout == LAYOUT_STRATA ? gw * gh : size_t( w ) * h;
    return px * d.comps * ( storage == RG_EXT_STORAGE_FP16 && IsHalfPlane( plane ) ? 2 : 4 );
}
void CheckSize( uint32_t w, uint32_t h, uint32_t storageMode )
{
    RG_REQUIRE( storageMode == RG_EXT_STORAGE_FP32 || storageMode == RG_EXT_STORAGE_FP16, RG_RESULT_WRONG_FUNCTION_ARGUMENT,
                "unknown storage mode" );
    RG_REQUIRE( w >= 2 && h >= 1 && ( w % 2 ) == 0, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render width must be even and >= 2" );
    RG_REQUIRE( uint64_t( w ) * h <= ( 1ull << 26 ), RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render size too large" );
}
} // namespace

size_t FramebufferSet::BytesFor( uint32_t w, uint32_t h, uint32_t storageMode, bool persistentOnly )
{
    size_t t = 0;
    for( int p = 0; p < RG_EXT_PLANE_COUNT; p++ )
    {
        if( !persistentOnly || IsPersistentPlane( p ) )
        {
            t += PlaneBytesFor( p, w, h, storageMode ) * ( g_planes[ p ].prev ? 2 : 1 );
        }
    }
    return t;
}

void TransientPlanes::Create( uint32_t maxW, uint32_t maxH, uint32_t storageMode )
{
    CheckSize( maxW, maxH, storageMode );
    capW    = maxW;
    capH    = maxH;
    storage = storageMode;
    for( int p = 0; p < RG_EXT_PLANE_COUNT; p++ )
    {
        planes[ p ].Release();
        if( !IsPersistentPlane( p ) )
        {
            const size_t bytes = PlaneBytesFor( p, maxW, maxH, storageMode );
            planes[ p ].Reserve( bytes );
            RG_HIP_CHECK( hipMemset( planes[ p ].ptr, 0, bytes ) );
        }
    }
}

size_t FramebufferSet::PlaneBytes( int plane ) const
{
    const PlaneDesc& d = g_planes[ plane ];
    const size_t     px = d.layout == LAYOUT_STRATA ? size_t( gw ) * gh : size_t( width ) * height;
    return px * d.comps * ( Half() && IsHalfPlane( plane ) ? 2 : 4 );
}


void FramebufferSet::Create( uint32_t w, uint32_t h, uint32_t storageMode, TransientPlanes* transient )
{
    CheckSize( w, h, storageMode );
    if( transient )
    {
        RG_REQUIRE( transient->storage == storageMode, RG_RESULT_WRONG_FUNCTION_ARGUMENT, "shared planes have another storage mode" );
        for( int p = 0; p < RG_EXT_PLANE_COUNT; p++ )
        {
            RG_REQUIRE( IsPersistentPlane( p ) || PlaneBytesFor( p, w, h, storageMode ) <= transient->planes[ p ].bytes,
                        RG_RESULT_WRONG_FUNCTION_ARGUMENT, "render size exceeds the shared planes' capacity" );
        }
    }
    storage = storageMode;
    shared  = transient;
    width  = w;
    height = h;
    gw     = ( w + 1 ) / 3 > 1 ? ( w + 1 ) / 3 : 1;
    gh     = ( h + 1 ) / 3 > 1 ? ( h + 1 ) / 3 : 1;
    parity = 0;
    for( int p = 0; p < RG_EXT_PLANE_COUNT; p++ )
    {
        const size_t bytes = PlaneBytes( p );
        for( int k = 0; k < ( g_planes[ p ].prev ? 2 : 1 ); k++ )
        {
            planes[ p ][ k ].Release();
            if( shared && !IsPersistentPlane( p ) )
            {
                continue;
            }
            planes[ p ][ k ].Reserve( bytes );
            RG_HIP_CHECK( hipMemset( planes[ p ][ k ].ptr, 0, bytes ) );
        }
    }
}

void* FramebufferSet::Cur( int plane )
{
    if( shared && !IsPersistentPlane( plane ) )
    {
        return shared->planes[ plane ].ptr;
    }
    return planes[ plane ][ g_planes[ plane ].prev ? parity : 0 ].ptr;
}

void* FramebufferSet::Prev( int plane )
{
    if( shared && !IsPersistentPlane( plane ) )
    {
        return shared->planes[ plane ].ptr;
    }
    return planes[ plane ][ g_planes[ plane ].prev ? ( parity ^ 1 ) : 0 ].ptr;
}

FbPtrs FramebufferSet::Ptrs()
{
    FbPtrs p;
    for( int i = 0; i < RG_EXT_PLANE_COUNT; i++ )
    {
        p.cur[ i ]  = Cur( i );
        p.prev[ i ] = Prev( i );
    }
    p.half = Half() ? 1u : 0u;
    return p;
}

size_t FramebufferSet::TotalBytes() const
{
    return BytesFor( width, height, storage, shared != nullptr );
}

} // namespace rg
