/* A C host of the RTGL1 API: the frame loop of the reference's own test
 * application (Tests/RtglTest.cpp:439-703) without a window.
 *
 *   - textures are provided outside a frame (a 1x1 cubemap, an albedo
 *     texture for the ground);
 *   - every frame: rgStartFrame, the ground's primitives in a shuffled order
 *     (upload order must not change the image, RtglTest.cpp:510-511), a
 *     moving box (RtglTest.cpp:534-565), a half-transparent quad that goes
 *     to the rasteriser (RtglTest.cpp:568-596), the sun (RtglTest.cpp:
 *     615-627) and rgDrawFrame with sky and render-resolution parameters
 *     and rayCullMaskWorld = WORLD_0 (RtglTest.cpp:648-699);
 *   - the protocol's error codes (RTGL1.cpp:46-71, VulkanDevice.cpp:828-847).
 *
 * Every RgResult is checked.  After the last frame the FINAL plane is read
 * back (rgExtGetFramebuffers, the build's readback extension) and written
 * as raw float32 RGBA to the output file.
 *
 * usage: rtgl_host OUTPUT FRAMES WIDTH HEIGHT */
#include <RTGL1/RTGL1.h>
#include <RTGL1/RTGL1_ext.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int g_failures = 0;

#define CHECK( call )                                                                                 \
    do                                                                                                \
    {                                                                                                 \
        RgResult r_ = ( call );                                                                       \
        if( r_ != RG_RESULT_SUCCESS )                                                                 \
        {                                                                                             \
            fprintf( stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #call, ( int )r_,          \
                     rgUtilGetResultDescription( r_ ) );                                              \
            exit( 1 );                                                                                \
        }                                                                                             \
    } while( 0 )

#define EXPECT( call, code )                                                                          \
    do                                                                                                \
    {                                                                                                 \
        RgResult r_ = ( call );                                                                       \
        if( r_ != ( code ) )                                                                          \
        {                                                                                             \
            fprintf( stderr, "%s:%d: %s -> %d, expected %d\n", __FILE__, __LINE__, #call, ( int )r_,  \
                     ( int )( code ) );                                                               \
            g_failures++;                                                                             \
        }                                                                                             \
    } while( 0 )

/* The scene, shared with the Python mirror of this program (the same
 * numbers, written there as float32). */
enum
{
    GROUND_TILES = 4,
};

/* ground: a 16 x 16 square at y = 0 as GROUND_TILES tiles of two triangles
 * each, one primitive per tile, indexed */
static void GroundTile( int tile, RgPrimitiveVertex v[ 4 ], uint32_t idx[ 6 ] )
{
    const float x0 = -8.0f + 8.0f * ( float )( tile % 2 ), z0 = -8.0f + 8.0f * ( float )( tile / 2 );
    const float xs[ 4 ] = { x0, x0 + 8.0f, x0 + 8.0f, x0 };
    const float zs[ 4 ] = { z0, z0, z0 + 8.0f, z0 + 8.0f };
    memset( v, 0, 4 * sizeof( RgPrimitiveVertex ) );
    for( int i = 0; i < 4; i++ )
    {
        v[ i ].position[ 0 ] = xs[ i ];
        v[ i ].position[ 2 ] = zs[ i ];
        v[ i ].normal[ 1 ]   = 1.0f;
        v[ i ].texCoord[ 0 ] = xs[ i ] * 0.25f;
        v[ i ].texCoord[ 1 ] = zs[ i ] * 0.25f;
        v[ i ].color         = 0xFFFFFFFFu;
    }
    /* counter-clockwise seen from above */
    const uint32_t t[ 6 ] = { 0, 2, 1, 0, 3, 2 };
    memcpy( idx, t, sizeof( t ) );
}

/* unit box centred at the origin: 12 triangles, 36 vertices, not indexed */
static void Box( RgPrimitiveVertex v[ 36 ] )
{
    static const float faceN[ 6 ][ 3 ] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
    memset( v, 0, 36 * sizeof( RgPrimitiveVertex ) );
    int k = 0;
    for( int f = 0; f < 6; f++ )
    {
        const float* n = faceN[ f ];
        /* two tangent axes with t x b = n, so the corners run counter-clockwise */
        float t[ 3 ], b[ 3 ];
        const int a = f / 2; /* axis of n */
        const int u = ( a + 1 ) % 3, w = ( a + 2 ) % 3;
        t[ 0 ] = t[ 1 ] = t[ 2 ] = 0.0f;
        b[ 0 ] = b[ 1 ] = b[ 2 ] = 0.0f;
        t[ u ] = 1.0f;
        b[ w ] = n[ a ];
        float c[ 4 ][ 3 ];
        const float su[ 4 ] = { -0.5f, 0.5f, 0.5f, -0.5f }, sw[ 4 ] = { -0.5f, -0.5f, 0.5f, 0.5f };
        for( int i = 0; i < 4; i++ )
        {
            for( int j = 0; j < 3; j++ )
            {
                c[ i ][ j ] = 0.5f * n[ j ] + su[ i ] * t[ j ] + sw[ i ] * b[ j ];
            }
        }
        const int order[ 6 ] = { 0, 1, 2, 0, 2, 3 };
        for( int i = 0; i < 6; i++, k++ )
        {
            memcpy( v[ k ].position, c[ order[ i ] ], sizeof( c[ 0 ] ) );
            memcpy( v[ k ].normal, n, sizeof( c[ 0 ] ) );
            v[ k ].color = 0xFFFFFFFFu;
        }
    }
}

/* the view matrix (column major, camera at (0, 6, 14) pitched down by
 * atan(3/4)); its entries are literals so that both hosts use the same floats */
static const float VIEW[ 16 ] = { 1.0f, 0.0f, 0.0f, 0.0f,  0.0f, 0.8f, 0.6f, 0.0f,
                                  0.0f, -0.6f, 0.8f, 0.0f, 0.0f, 3.6f, -14.8f, 1.0f };

static uint32_t g_lcg = 12345u;
static uint32_t NextRandom( void )
{
    g_lcg = g_lcg * 1664525u + 1013904223u;
    return g_lcg >> 8;
}

int main( int argc, char** argv )
{
    if( argc != 5 )
    {
        fprintf( stderr, "usage: %s OUTPUT FRAMES WIDTH HEIGHT\n", argv[ 0 ] );
        return 2;
    }
    const char*    outPath = argv[ 1 ];
    const int      frames  = atoi( argv[ 2 ] );
    const uint32_t width = ( uint32_t )atoi( argv[ 3 ] ), height = ( uint32_t )atoi( argv[ 4 ] );

    RgInstanceCreateInfo ci;
    memset( &ci, 0, sizeof( ci ) );
    ci.pAppName      = "rtgl_host";
    ci.pAppGUID      = "c4e3b8a0-0000-4000-8000-rtglhost0001";
    ci.worldUp.data[ 1 ] = 1.0f;
    ci.worldForward.data[ 2 ] = 1.0f;
    ci.worldScale        = 1.0f;

    /* protocol errors before an instance exists */
    EXPECT( rgCreateInstance( &ci, NULL ), RG_RESULT_WRONG_FUNCTION_ARGUMENT );

    RgInstance instance = NULL;
    CHECK( rgCreateInstance( &ci, &instance ) );
    {
        RgInstance second = NULL;
        EXPECT( rgCreateInstance( &ci, &second ), RG_RESULT_ALREADY_INITIALIZED );
        EXPECT( rgStartFrame( ( RgInstance )( ( char* )instance + 1 ), NULL ), RG_RESULT_WRONG_INSTANCE );
    }

    /* resources provided outside a frame (RtglTest.cpp:445-477) */
    {
        const uint32_t        white = 0xFFFFFFFFu;
        RgOriginalCubemapInfo sky;
        memset( &sky, 0, sizeof( sky ) );
        sky.pTextureName     = "_external_/cubemap/0";
        sky.pPixelsPositiveX = &white;
        sky.pPixelsNegativeX = &white;
        sky.pPixelsPositiveY = &white;
        sky.pPixelsNegativeY = &white;
        sky.pPixelsPositiveZ = &white;
        sky.pPixelsNegativeZ = &white;
        sky.sideSize         = 1;
        CHECK( rgProvideOriginalCubemapTexture( instance, &sky ) );

        /* a 4 x 4 checker of two greys (sRGB) */
        uint32_t px[ 16 ];
        for( int i = 0; i < 16; i++ )
        {
            px[ i ] = ( ( i % 4 + i / 4 ) % 2 ) ? rgUtilPackColorByte4D( 200, 200, 200, 255 )
                                                : rgUtilPackColorByte4D( 90, 110, 90, 255 );
        }
        RgOriginalTextureInfo tex;
        memset( &tex, 0, sizeof( tex ) );
        tex.pTextureName = "ground";
        tex.pPixels      = px;
        tex.size.width   = 4;
        tex.size.height  = 4;
        CHECK( rgProvideOriginalTexture( instance, &tex ) );
    }

    EXPECT( rgDrawFrame( instance, NULL ), RG_RESULT_WRONG_FUNCTION_ARGUMENT );
    {
        RgDrawFrameInfo early;
        memset( &early, 0, sizeof( early ) );
        EXPECT( rgDrawFrame( instance, &early ), RG_RESULT_FRAME_WASNT_STARTED );
    }

    RgPrimitiveVertex tileV[ GROUND_TILES ][ 4 ];
    uint32_t          tileI[ GROUND_TILES ][ 6 ];
    for( int t = 0; t < GROUND_TILES; t++ )
    {
        GroundTile( t, tileV[ t ], tileI[ t ] );
    }
    RgPrimitiveVertex box[ 36 ];
    Box( box );
    RgPrimitiveVertex quad[ 6 ];
    memcpy( quad, box, sizeof( quad ) ); /* one face of the box as a flat quad */

    for( int frame = 0; frame < frames; frame++ )
    {
        RgStartFrameInfo sfi;
        memset( &sfi, 0, sizeof( sfi ) );
        sfi.pMapName = "untitled";
        CHECK( rgStartFrame( instance, &sfi ) );
        if( frame == 0 )
        {
            EXPECT( rgStartFrame( instance, &sfi ), RG_RESULT_FRAME_WASNT_ENDED );
        }

        /* the ground, its tiles in a random order */
        {
            RgMeshInfo mesh;
            memset( &mesh, 0, sizeof( mesh ) );
            mesh.uniqueObjectID = 1;
            mesh.pMeshName      = "ground";
            mesh.transform.matrix[ 0 ][ 0 ] = mesh.transform.matrix[ 1 ][ 1 ] = mesh.transform.matrix[ 2 ][ 2 ] = 1.0f;
            mesh.isExportable   = 1;
            int order[ GROUND_TILES ];
            for( int i = 0; i < GROUND_TILES; i++ )
            {
                order[ i ] = i;
            }
            for( int i = GROUND_TILES - 1; i > 0; i-- )
            {
                const int j = ( int )( NextRandom() % ( uint32_t )( i + 1 ) );
                const int s = order[ i ];
                order[ i ]  = order[ j ];
                order[ j ]  = s;
            }
            for( int i = 0; i < GROUND_TILES; i++ )
            {
                const int           t = order[ i ];
                RgMeshPrimitiveInfo prim;
                memset( &prim, 0, sizeof( prim ) );
                prim.primitiveIndexInMesh = ( uint32_t )t;
                prim.pVertices            = tileV[ t ];
                prim.vertexCount          = 4;
                prim.pIndices             = tileI[ t ];
                prim.indexCount           = 6;
                prim.pTextureName         = "ground";
                prim.color                = 0xFFFFFFFFu;
                CHECK( rgUploadMeshPrimitive( instance, &mesh, &prim ) );
            }
        }

        /* the moving box (RtglTest.cpp:534-565) */
        {
            RgMeshInfo mesh;
            memset( &mesh, 0, sizeof( mesh ) );
            mesh.uniqueObjectID = 10;
            mesh.pMeshName      = "test";
            mesh.transform.matrix[ 0 ][ 0 ] = mesh.transform.matrix[ 1 ][ 1 ] = mesh.transform.matrix[ 2 ][ 2 ] = 1.0f;
            mesh.transform.matrix[ 0 ][ 3 ] = 5.0f - 0.05f * ( float )( ( frame + 30 ) % 200 );
            mesh.transform.matrix[ 1 ][ 3 ] = 1.0f;
            RgMeshPrimitiveInfo prim;
            memset( &prim, 0, sizeof( prim ) );
            prim.pVertices   = box;
            prim.vertexCount = 36;
            prim.color       = rgUtilPackColorByte4D( 128, 255, 128, 255 );
            CHECK( rgUploadMeshPrimitive( instance, &mesh, &prim ) );
        }

        /* a half-transparent quad: rasterised, not ray traced (RtglTest.cpp:568-596) */
        {
            RgMeshInfo mesh;
            memset( &mesh, 0, sizeof( mesh ) );
            mesh.uniqueObjectID = 12;
            mesh.pMeshName      = "test_raster";
            mesh.transform.matrix[ 0 ][ 0 ] = mesh.transform.matrix[ 1 ][ 1 ] = mesh.transform.matrix[ 2 ][ 2 ] = 1.0f;
            mesh.transform.matrix[ 0 ][ 3 ] = -0.5f;
            mesh.transform.matrix[ 1 ][ 3 ] = 1.0f;
            mesh.transform.matrix[ 2 ][ 3 ] = 1.0f;
            RgMeshPrimitiveInfo prim;
            memset( &prim, 0, sizeof( prim ) );
            prim.pVertices   = quad;
            prim.vertexCount = 6;
            prim.color       = rgUtilPackColorByte4D( 255, 128, 128, 128 );
            CHECK( rgUploadMeshPrimitive( instance, &mesh, &prim ) );
        }
        EXPECT( rgUploadMeshPrimitive( instance, NULL, NULL ), RG_RESULT_WRONG_FUNCTION_ARGUMENT );

        /* the sun (RtglTest.cpp:615-627) */
        {
            RgDirectionalLightUploadInfo sun;
            memset( &sun, 0, sizeof( sun ) );
            sun.uniqueID               = 0;
            sun.isExportable           = 1;
            sun.color                  = rgUtilPackColorByte4D( 255, 255, 255, 255 );
            sun.intensity              = 7000.0f;
            sun.direction.data[ 0 ]    = -1.0f;
            sun.direction.data[ 1 ]    = -8.0f;
            sun.direction.data[ 2 ]    = -1.0f;
            sun.angularDiameterDegrees = 0.5f;
            CHECK( rgUploadDirectionalLight( instance, &sun ) );
        }

        /* submit (RtglTest.cpp:648-699), at a custom render size */
        {
            RgDrawFrameSkyParams sky;
            memset( &sky, 0, sizeof( sky ) );
            sky.sType                   = RG_STRUCTURE_TYPE_SKY;
            sky.skyType                 = RG_SKY_TYPE_COLOR;
            sky.skyColorDefault.data[ 0 ] = 0.71f;
            sky.skyColorDefault.data[ 1 ] = 0.88f;
            sky.skyColorDefault.data[ 2 ] = 1.0f;
            sky.skyColorMultiplier      = 1000.0f;
            sky.skyColorSaturation      = 1.0f;
            sky.pSkyCubemapTextureName  = "_external_/cubemap/0";

            RgDrawFrameRenderResolutionParams res;
            memset( &res, 0, sizeof( res ) );
            res.sType                   = RG_STRUCTURE_TYPE_RENDER_RESOLUTION;
            res.pNext                   = &sky;
            res.upscaleTechnique        = RG_RENDER_UPSCALE_TECHNIQUE_LINEAR;
            res.resolutionMode          = RG_RENDER_RESOLUTION_MODE_CUSTOM;
            res.customRenderSize.width  = width;
            res.customRenderSize.height = height;

            RgDrawFrameInfo fi;
            memset( &fi, 0, sizeof( fi ) );
            memcpy( fi.view, VIEW, sizeof( VIEW ) );
            fi.fovYRadians      = 1.3089969f; /* 75 degrees */
            fi.cameraNear       = 0.1f;
            fi.cameraFar        = 10000.0f;
            fi.rayLength        = 10000.0f;
            fi.rayCullMaskWorld = RG_DRAW_FRAME_RAY_CULL_WORLD_0_BIT;
            fi.currentTime      = ( double )( frame + 1 ) / 60.0;
            fi.pParams          = &res;
            CHECK( rgDrawFrame( instance, &fi ) );
        }
    }

    /* readback of the last frame's FINAL plane */
    RgExtFramebuffers fb = NULL;
    CHECK( rgExtGetFramebuffers( instance, &fb ) );
    void*  d_plane = NULL;
    size_t bytes   = 0;
    CHECK( rgExtFramebuffersGetPlane( fb, RG_EXT_PLANE_FINAL, 0, &d_plane, &bytes ) );
    if( bytes != ( size_t )width * height * 16 )
    {
        fprintf( stderr, "FINAL is %zu bytes, expected %zu\n", bytes, ( size_t )width * height * 16 );
        return 1;
    }
    float* host = ( float* )malloc( bytes );
    CHECK( rgExtCopyDeviceToHost( host, d_plane, bytes ) );
    FILE* f = fopen( outPath, "wb" );
    if( !f || fwrite( host, 1, bytes, f ) != bytes )
    {
        fprintf( stderr, "cannot write %s\n", outPath );
        return 1;
    }
    fclose( f );
    free( host );

    CHECK( rgDestroyInstance( instance ) );
    EXPECT( rgDestroyInstance( instance ), RG_RESULT_WRONG_INSTANCE );
    if( g_failures )
    {
        fprintf( stderr, "%d protocol checks failed\n", g_failures );
        return 1;
    }
    printf( "rtgl_host ok: %d frames at %ux%u\n", frames, width, height );
    return 0;
}
