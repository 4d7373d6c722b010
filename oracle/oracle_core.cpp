// CPU ORACLE - TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Codecs, RNG, the LBVH restatement, traversal and eye adaptation.
// Built with -ffp-contract=off so FP32 expressions round exactly as written.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace
{

template< typename To, typename From >
To bitcast( From f )
{
    static_assert( sizeof( To ) == sizeof( From ), "size" );
    To t;
    std::memcpy( &t, &f, sizeof( To ) );
    return t;
}

template< typename F >
void ParallelFor( uint32_t count, int threads, F&& f )
{
    if( threads <= 1 || count < 1024 )
    {
        for( uint32_t i = 0; i < count; i++ )
        {
            f( i );
        }
        return;
    }
    std::atomic< uint32_t > next{ 0 };
    std::vector< std::thread > pool;
    const uint32_t chunk = 256;
    for( int t = 0; t < threads; t++ )
    {
        pool.emplace_back( [ & ] {
            while( true )
            {
                uint32_t b = next.fetch_add( chunk );
                if( b >= count )
                {
                    break;
                }
                uint32_t e = std::min( count, b + chunk );
                for( uint32_t i = b; i < e; i++ )
                {
                    f( i );
                }
            }
        } );
    }
    for( auto& th : pool )
    {
        th.join();
    }
}

struct F3
{
    float x, y, z;
};
inline F3 sub( F3 a, F3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
inline float comp( F3 a, int i ) { return i == 0 ? a.x : ( i == 1 ? a.y : a.z ); }

} // namespace

// =====================================================================
// Portable log2 / exp2 (same operation sequence as the kernels).
// log2 via the atanh series around 1; exp2 via range reduction and the
// Taylor series of e^y (see raytracedgl1 device math for the derivation).
// =====================================================================
extern "C" float or_log2f( float x )
{
    uint32_t bits = bitcast< uint32_t >( x );
    int      e    = int( ( bits >> 23 ) & 0xFFu );
    if( e == 0 )
    {
        float xs = x * 4294967296.0f;
        bits     = bitcast< uint32_t >( xs );
        e        = int( ( bits >> 23 ) & 0xFFu ) - 32;
    }
    e -= 127;
    float m = bitcast< float >( ( bits & 0x007FFFFFu ) | 0x3F800000u );
    if( m > 1.41421356f )
    {
        m = m * 0.5f;
        e += 1;
    }
    const float t  = ( m - 1.0f ) / ( m + 1.0f );
    const float t2 = t * t;
    const float s  = t2 * ( 0.333333333f + t2 * ( 0.2f + t2 * ( 0.142857143f + t2 * ( 0.111111111f + t2 * 0.0909090909f ) ) ) );
    return float( e ) + 2.88539008f * ( t + t * s );
}

extern "C" float or_exp2f( float x )
{
    if( !( x > -126.0f ) )
    {
        return x != x ? x : 0.0f;
    }
    if( x >= 128.0f )
    {
        return INFINITY;
    }
    const float n = std::floor( x + 0.5f );
    const float f = x - n;
    const float y = f * 0.693147181f;
    const float p = 1.0f + y * ( 1.0f + y * ( 0.5f + y * ( 0.166666667f + y * ( 0.0416666667f + y * ( 0.00833333333f + y * ( 0.00138888889f + y * ( 0.000198412698f + y * 0.0000248015873f ) ) ) ) ) ) );
    const int ni = int( n );
    const int h  = ni / 2;
    const float s1 = bitcast< float >( uint32_t( h + 127 ) << 23 );
    const float s2 = bitcast< float >( uint32_t( ni - h + 127 ) << 23 );
    return ( p * s1 ) * s2;
}

// =====================================================================
// RNG: Source/Shaders/Random.h:243-300.
// =====================================================================
extern "C" uint32_t or_wellons_low_bias32( uint32_t x )
{
    // Random.h:244-252
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

extern "C" void or_murmur_hash33( uint32_t x, uint32_t y, uint32_t z, uint32_t out[ 3 ] )
{
    // Random.h:277-284: src *= M; src ^= src >> 24; src *= M; then the
    // three-lane mix of h.
    const uint32_t M      = 0x5bd1e995u;
    uint32_t       src[3] = { x, y, z };
    for( auto& s : src )
    {
        s *= M;
        s ^= s >> 24u;
        s *= M;
    }
    uint32_t h[ 3 ] = { 1190494759u, 2147483647u, 3559788179u };
    for( auto& v : h )
    {
        v *= M;
    }
    for( auto& v : h ) v ^= src[ 0 ];
    for( auto& v : h ) v *= M;
    for( auto& v : h ) v ^= src[ 1 ];
    for( auto& v : h ) v *= M;
    for( auto& v : h ) v ^= src[ 2 ];
    for( auto& v : h ) v ^= v >> 13u;
    for( auto& v : h ) v *= M;
    for( auto& v : h ) v ^= v >> 15u;
    out[ 0 ] = h[ 0 ];
    out[ 1 ] = h[ 1 ];
    out[ 2 ] = h[ 2 ];
}

extern "C" uint32_t or_random_seed( int px, int py, uint32_t frameIndex )
{
    // Random.h:286-300 with packRandomSeed (:205-211), 128 layers of 128x128.
    uint32_t h[ 3 ];
    or_murmur_hash33( uint32_t( px ), uint32_t( py ), frameIndex, h );
    const uint32_t ox = h[ 0 ] % 128u, oy = h[ 1 ] % 128u, layer = h[ 2 ] % 128u;
    return ( layer << 14 ) | ( oy << 7 ) | ox;
}

// =====================================================================
// Codecs: Source/Shaders/Utils.h:108-207.
// =====================================================================
extern "C" uint32_t or_encode_normal( const double n[ 3 ] )
{
    // Utils.h:108-124: 16-bit polar angle phi = acos(z), 16-bit azimuth.
    const double   pi     = 3.14159265358979323846;
    const uint32_t Nphi   = 1u << 16;
    const uint32_t Ntheta = 1u << 16;
    double         phi    = std::acos( std::max( -1.0, std::min( 1.0, n[ 2 ] ) ) );
    double         theta  = std::atan2( n[ 1 ], n[ 0 ] );
    theta                 = theta < 0 ? theta + 2 * pi : theta;
    uint32_t j            = uint32_t( std::llround( phi * ( Nphi - 1 ) / pi ) );
    uint32_t k            = uint32_t( std::llround( theta * Ntheta / ( 2 * pi ) ) ) % Ntheta;
    return ( j << 16 ) | k;
}

extern "C" void or_decode_normal( uint32_t packed, double n[ 3 ] )
{
    // Utils.h:126-140
    const double   pi     = 3.14159265358979323846;
    const uint32_t Nphi   = 1u << 16;
    const uint32_t Ntheta = 1u << 16;
    const uint32_t j      = packed >> 16;
    const uint32_t k      = packed & 0xFFFFu;
    const double   phi    = j * pi / ( Nphi - 1 );
    const double   theta  = k * 2 * pi / Ntheta;
    n[ 0 ]                = std::sin( phi ) * std::cos( theta );
    n[ 1 ]                = std::sin( phi ) * std::sin( theta );
    n[ 2 ]                = std::cos( phi );
}

extern "C" uint32_t or_encode_e5b9g9r9( const double rgbIn[ 3 ] )
{
    // Utils.h:162-192, the EXT_texture_shared_exponent encoding
    // (N = 9 mantissa bits, B = 15 bias, max 65408).
    const int N = 9, B = 15, Np2 = 1 << N;
    double    rgb[ 3 ];
    for( int i = 0; i < 3; i++ )
    {
        rgb[ i ] = std::min( std::max( rgbIn[ i ], 0.0 ), 65408.0 );
    }
    const double maxc = std::max( rgb[ 0 ], std::max( rgb[ 1 ], rgb[ 2 ] ) );
    if( maxc == 0.0 )
    {
        return 0;
    }
    const int expP = std::max( -B - 1, int( std::floor( std::log2( maxc ) ) ) ) + 1 + B;
    const int maxS = int( std::nearbyint( maxc * std::exp2( -( expP - B - N ) ) ) );
    const int exp  = maxS != Np2 ? expP : expP + 1;
    const double s = std::exp2( -( exp - B - N ) );
    uint32_t     q[ 3 ];
    for( int i = 0; i < 3; i++ )
    {
        q[ i ] = uint32_t( std::nearbyint( rgb[ i ] * s ) );
    }
    return ( uint32_t( exp ) << 27 ) | ( q[ 2 ] << 18 ) | ( q[ 1 ] << 9 ) | q[ 0 ];
}

extern "C" void or_decode_e5b9g9r9( uint32_t p, double rgb[ 3 ] )
{
    // Utils.h:194-207
    const int    exp = int( p >> 27 );
    const double s   = std::exp2( double( exp - 15 - 9 ) );
    rgb[ 0 ]         = s * double( p & 511u );
    rgb[ 1 ]         = s * double( ( p >> 9 ) & 511u );
    rgb[ 2 ]         = s * double( ( p >> 18 ) & 511u );
}

extern "C" void or_checkerboard_from_regular( int x, int y, int width, int* cx, int* cy )
{
    // ShaderCommonGLSLFunc.h:321-331
    const int sep = width / 2;
    const int odd = ( x + y % 2 ) % 2;
    *cx           = odd * sep + x / 2;
    *cy           = y;
}

extern "C" void or_checkerboard_to_regular( int cx, int cy, int width, int* x, int* y )
{
    // ShaderCommonGLSLFunc.h:333-344
    const int sep = width / 2;
    const int odd = cx >= sep ? 1 : 0;
    *x            = ( cx - odd * sep ) * 2 + ( odd + cy ) % 2;
    *y            = cy;
}

// =====================================================================
// LBVH.  The reference builds acceleration structures in the Vulkan driver
// (Source/ASBuilder.cpp:141-206); this restates the algorithms the GPU build
// implements:
//   1. FP64 centroids ((v0 + v1) + v2) / 3 and their bounds;
//   2. 21-bit-per-axis quantisation, 63-bit Morton codes (x in the highest
//      bit of each triple);
//   3. stable sort of (code, index);
//   4. Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
//      and k-d Trees", HPG 2012 (delta with index tie-break);
//   5. bottom-up AABB union;
//   6. 8-wide collapse: open the child of largest surface area among the
//      internal children holding more than leafMax triangles (first one on
//      ties), new children take the opened child's slot and the next one;
//      8-wide nodes are numbered by increasing binary node index.
// =====================================================================
struct OrBvh
{
    uint32_t                n = 0, leafMax = 3;
    double                  lo[ 3 ], hi[ 3 ];
    std::vector< uint64_t > codes;       // input order
    std::vector< uint32_t > sortedIdx;   // n
    std::vector< uint64_t > sortedKeys;  // n
    std::vector< uint32_t > children;    // 2(n-1)
    std::vector< uint32_t > rf, rl;      // n-1
    std::vector< float >    bounds;      // 6 (2n-1)
    std::vector< uint8_t >  nodes;       // nodeCount * 256
    std::vector< float >    sortedTris;  // n * 12 (3 float4 per tri, w of the first = id bits)
    uint32_t                nodeCount = 0;
};

namespace
{
constexpr uint32_t CHILD_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t LEAF_BIT    = 0x80000000u;

uint64_t Expand21( uint32_t v )
{
    uint64_t r = 0;
    for( int b = 0; b < 21; b++ )
    {
        r |= uint64_t( ( v >> b ) & 1u ) << ( 3 * b );
    }
    return r;
}

int Clz64( uint64_t x )
{
    return x == 0 ? 64 : __builtin_clzll( x );
}
int Clz32( uint32_t x )
{
    return x == 0 ? 32 : __builtin_clz( x );
}

int Delta( const std::vector< uint64_t >& k, int64_t n, int64_t i, int64_t j )
{
    if( j < 0 || j >= n )
    {
        return -1;
    }
    if( k[ i ] == k[ j ] )
    {
        return 64 + Clz32( uint32_t( i ) ^ uint32_t( j ) );
    }
    return Clz64( k[ i ] ^ k[ j ] );
}

float HalfArea( const float* b )
{
    const float dx = b[ 3 ] - b[ 0 ], dy = b[ 4 ] - b[ 1 ], dz = b[ 5 ] - b[ 2 ];
    return dx * dy + dy * dz + dz * dx;
}

struct Node8
{
    float    lo_x[ 8 ], hi_x[ 8 ], lo_y[ 8 ], hi_y[ 8 ], lo_z[ 8 ], hi_z[ 8 ];
    uint32_t child[ 8 ];
    uint32_t pad[ 8 ];
};
static_assert( sizeof( Node8 ) == 256, "node" );
} // namespace

extern "C" OrBvh* or_bvh_build( const float* tris, uint32_t n, uint32_t leafMax )
{
    if( n == 0 || leafMax < 1 || leafMax > 8 )
    {
        return nullptr;
    }
    OrBvh* b   = new OrBvh();
    b->n       = n;
    b->leafMax = leafMax;

    // 1. centroids and bounds
    std::vector< double > cen( size_t( n ) * 3 );
    for( int k = 0; k < 3; k++ )
    {
        b->lo[ k ] = DBL_MAX;
        b->hi[ k ] = -DBL_MAX;
    }
    for( uint32_t i = 0; i < n; i++ )
    {
        const float* t = tris + size_t( i ) * 9;
        for( int k = 0; k < 3; k++ )
        {
            double c          = ( ( double( t[ k ] ) + double( t[ 3 + k ] ) ) + double( t[ 6 + k ] ) ) / 3.0;
            cen[ i * 3 + k ]  = c;
            b->lo[ k ]        = std::min( b->lo[ k ], c );
            b->hi[ k ]        = std::max( b->hi[ k ], c );
        }
    }
    // 2. Morton codes
    b->codes.resize( n );
    for( uint32_t i = 0; i < n; i++ )
    {
        uint32_t q[ 3 ];
        for( int k = 0; k < 3; k++ )
        {
            const double ext = b->hi[ k ] - b->lo[ k ];
            if( !( ext > 0.0 ) )
            {
                q[ k ] = 0;
                continue;
            }
            const double f = std::floor( ( ( cen[ i * 3 + k ] - b->lo[ k ] ) / ext ) * 2097152.0 );
            q[ k ]         = f >= 2097151.0 ? 2097151u : ( f <= 0.0 ? 0u : uint32_t( f ) );
        }
        b->codes[ i ] = ( Expand21( q[ 0 ] ) << 2 ) | ( Expand21( q[ 1 ] ) << 1 ) | Expand21( q[ 2 ] );
    }
    // 3. stable sort
    b->sortedIdx.resize( n );
    for( uint32_t i = 0; i < n; i++ )
    {
        b->sortedIdx[ i ] = i;
    }
    std::stable_sort( b->sortedIdx.begin(), b->sortedIdx.end(), [ & ]( uint32_t x, uint32_t y ) { return b->codes[ x ] < b->codes[ y ]; } );
    b->sortedKeys.resize( n );
    for( uint32_t i = 0; i < n; i++ )
    {
        b->sortedKeys[ i ] = b->codes[ b->sortedIdx[ i ] ];
    }
    // 4. Karras hierarchy
    const int64_t N = n;
    const uint32_t internal = n - 1;
    b->children.assign( size_t( internal ) * 2, 0 );
    b->rf.assign( internal, 0 );
    b->rl.assign( internal, 0 );
    const auto& K = b->sortedKeys;
    for( int64_t i = 0; i < N - 1; i++ )
    {
        const int d    = ( Delta( K, N, i, i + 1 ) - Delta( K, N, i, i - 1 ) ) >= 0 ? 1 : -1;
        const int dmin = Delta( K, N, i, i - d );
        int64_t   lmax = 2;
        while( Delta( K, N, i, i + lmax * d ) > dmin )
        {
            lmax *= 2;
        }
        int64_t l = 0;
        for( int64_t t = lmax / 2; t >= 1; t /= 2 )
        {
            if( Delta( K, N, i, i + ( l + t ) * d ) > dmin )
            {
                l += t;
            }
        }
        const int64_t j     = i + l * d;
        const int     dnode = Delta( K, N, i, j );
        int64_t       s = 0, t = l;
        do
        {
            t = ( t + 1 ) / 2;
            if( Delta( K, N, i, i + ( s + t ) * d ) > dnode )
            {
                s += t;
            }
        } while( t > 1 );
        const int64_t gamma = i + s * d + std::min( d, 0 );
        const int64_t lo = std::min( i, j ), hi = std::max( i, j );
        b->children[ i * 2 + 0 ] = uint32_t( lo == gamma ? ( N - 1 ) + gamma : gamma );
        b->children[ i * 2 + 1 ] = uint32_t( hi == gamma + 1 ? ( N - 1 ) + gamma + 1 : gamma + 1 );
        b->rf[ i ]               = uint32_t( lo );
        b->rl[ i ]               = uint32_t( hi );
    }
    // 5. bounds: leaves, then internal nodes children-first (post-order)
    b->bounds.assign( size_t( 2 * n - 1 ) * 6, 0.0f );
    for( uint32_t k = 0; k < n; k++ )
    {
        const float* t  = tris + size_t( b->sortedIdx[ k ] ) * 9;
        float*       bb = &b->bounds[ size_t( internal + k ) * 6 ];
        for( int a = 0; a < 3; a++ )
        {
            bb[ a ]     = std::fmin( std::fmin( t[ a ], t[ 3 + a ] ), t[ 6 + a ] );
            bb[ 3 + a ] = std::fmax( std::fmax( t[ a ], t[ 3 + a ] ), t[ 6 + a ] );
        }
    }
    if( internal > 0 )
    {
        std::vector< std::pair< uint32_t, bool > > st;
        st.push_back( { 0u, false } );
        while( !st.empty() )
        {
            auto [ node, visited ] = st.back();
            st.pop_back();
            if( node >= internal )
            {
                continue;
            }
            if( !visited )
            {
                st.push_back( { node, true } );
                st.push_back( { b->children[ node * 2 + 0 ], false } );
                st.push_back( { b->children[ node * 2 + 1 ], false } );
                continue;
            }
            const float* l  = &b->bounds[ size_t( b->children[ node * 2 + 0 ] ) * 6 ];
            const float* r  = &b->bounds[ size_t( b->children[ node * 2 + 1 ] ) * 6 ];
            float*       bb = &b->bounds[ size_t( node ) * 6 ];
            for( int a = 0; a < 3; a++ )
            {
                bb[ a ]     = std::fmin( l[ a ], r[ a ] );
                bb[ 3 + a ] = std::fmax( l[ 3 + a ], r[ 3 + a ] );
            }
        }
    }
    // 6. collapse
    auto count = [ & ]( uint32_t id ) -> uint32_t { return id >= internal ? 1u : b->rl[ id ] - b->rf[ id ] + 1u; };
    const uint32_t slots = internal > 0 ? internal : 1;
    std::vector< std::vector< uint32_t > > expansion( slots );
    std::vector< uint8_t >                 isNode8( slots, 0 );
    std::vector< uint32_t >                queue = { 0u };
    for( size_t qi = 0; qi < queue.size(); qi++ )
    {
        const uint32_t          bn = queue[ qi ];
        std::vector< uint32_t > ch;
        if( bn >= internal || count( bn ) <= leafMax )
        {
            ch = { bn };
        }
        else
        {
            ch = { b->children[ bn * 2 ], b->children[ bn * 2 + 1 ] };
            while( ch.size() < 8 )
            {
                int   best     = -1;
                float bestArea = -1.0f;
                for( int k = 0; k < int( ch.size() ); k++ )
                {
                    if( ch[ k ] < internal && count( ch[ k ] ) > leafMax )
                    {
                        const float a = HalfArea( &b->bounds[ size_t( ch[ k ] ) * 6 ] );
                        if( a > bestArea )
                        {
                            bestArea = a;
                            best     = k;
                        }
                    }
                }
                if( best < 0 )
                {
                    break;
                }
                const uint32_t c = ch[ best ];
                ch[ best ]       = b->children[ c * 2 ];
                ch.insert( ch.begin() + best + 1, b->children[ c * 2 + 1 ] );
            }
        }
        const uint32_t slot = bn < internal ? bn : 0;
        expansion[ slot ]   = ch;
        isNode8[ slot ]     = 1;
        if( ch.size() > 1 )
        {
            for( uint32_t c : ch )
            {
                if( c < internal && count( c ) > leafMax )
                {
                    queue.push_back( c );
                }
            }
        }
    }
    std::vector< uint32_t > index( slots, 0 );
    uint32_t                total = 0;
    for( uint32_t s = 0; s < slots; s++ )
    {
        index[ s ] = total;
        total += isNode8[ s ];
    }
    b->nodeCount = total;
    b->nodes.assign( size_t( total ) * sizeof( Node8 ), 0 );
    Node8* nodes = reinterpret_cast< Node8* >( b->nodes.data() );
    for( uint32_t s = 0; s < slots; s++ )
    {
        if( !isNode8[ s ] )
        {
            continue;
        }
        Node8&      nd = nodes[ index[ s ] ];
        const auto& ch = expansion[ s ];
        for( int k = 0; k < 8; k++ )
        {
            nd.pad[ k ] = 0;
            if( k >= int( ch.size() ) )
            {
                nd.lo_x[ k ] = nd.lo_y[ k ] = nd.lo_z[ k ] = INFINITY;
                nd.hi_x[ k ] = nd.hi_y[ k ] = nd.hi_z[ k ] = -INFINITY;
                nd.child[ k ] = CHILD_EMPTY;
                continue;
            }
            const uint32_t c  = ch[ k ];
            const float*   bb = &b->bounds[ size_t( c ) * 6 ];
            nd.lo_x[ k ]      = bb[ 0 ];
            nd.lo_y[ k ]      = bb[ 1 ];
            nd.lo_z[ k ]      = bb[ 2 ];
            nd.hi_x[ k ]      = bb[ 3 ];
            nd.hi_y[ k ]      = bb[ 4 ];
            nd.hi_z[ k ]      = bb[ 5 ];
            if( c >= internal )
            {
                nd.child[ k ] = LEAF_BIT | ( ( c - internal ) << 3 ) | 0u;
            }
            else if( count( c ) <= leafMax || !isNode8[ c ] || ch.size() == 1 )
            {
                nd.child[ k ] = LEAF_BIT | ( b->rf[ c ] << 3 ) | ( count( c ) - 1u );
            }
            else
            {
                nd.child[ k ] = index[ c ];
            }
        }
    }
    // sorted triangles for traversal
    b->sortedTris.assign( size_t( n ) * 12, 0.0f );
    for( uint32_t k = 0; k < n; k++ )
    {
        const uint32_t id = b->sortedIdx[ k ];
        const float*   t  = tris + size_t( id ) * 9;
        float*         o  = &b->sortedTris[ size_t( k ) * 12 ];
        o[ 0 ] = t[ 0 ]; o[ 1 ] = t[ 1 ]; o[ 2 ] = t[ 2 ]; o[ 3 ] = bitcast< float >( id );
        o[ 4 ] = t[ 3 ]; o[ 5 ] = t[ 4 ]; o[ 6 ] = t[ 5 ]; o[ 7 ] = 0.0f;
        o[ 8 ] = t[ 6 ]; o[ 9 ] = t[ 7 ]; o[ 10 ] = t[ 8 ]; o[ 11 ] = 0.0f;
    }
    return b;
}

extern "C" void or_bvh_free( OrBvh* b )
{
    delete b;
}
extern "C" uint32_t or_bvh_node_count( const OrBvh* b )
{
    return b ? b->nodeCount : 0;
}
extern "C" void or_bvh_scene_bounds( const OrBvh* b, double outMin[ 3 ], double outMax[ 3 ] )
{
    for( int k = 0; k < 3; k++ )
    {
        outMin[ k ] = b->lo[ k ];
        outMax[ k ] = b->hi[ k ];
    }
}
extern "C" void or_bvh_copy_codes( const OrBvh* b, uint64_t* out )
{
    std::memcpy( out, b->codes.data(), b->codes.size() * 8 );
}
extern "C" void or_bvh_copy_sorted_indices( const OrBvh* b, uint32_t* out )
{
    std::memcpy( out, b->sortedIdx.data(), b->sortedIdx.size() * 4 );
}
extern "C" void or_bvh_copy_binary( const OrBvh* b, uint32_t* children, float* bounds )
{
    if( children && !b->children.empty() )
    {
        std::memcpy( children, b->children.data(), b->children.size() * 4 );
    }
    if( bounds )
    {
        std::memcpy( bounds, b->bounds.data(), b->bounds.size() * 4 );
    }
}
extern "C" void or_bvh_copy_nodes( const OrBvh* b, void* out )
{
    std::memcpy( out, b->nodes.data(), b->nodes.size() );
}

// =====================================================================
// Traversal: the same watertight test (Woop, Benthin, Wald, JCGT 2013),
// robust slab test (Ize, JCGT 2013) and closest-hit rule as the kernel:
// a hit replaces the best one if t is smaller, or equal with a smaller id.
// Hit semantics follow traceRayEXT as used in RaygenCommon.h:217-359.
// =====================================================================
namespace
{
struct Pre
{
    F3    o, d, inv;
    int   kx, ky, kz;
    float Sx, Sy, Sz;
};

float Guard( float d )
{
    return std::fabs( d ) >= 1e-20f ? d : ( d < 0.0f ? -1e-20f : 1e-20f );
}

Pre MakePre( F3 o, F3 d )
{
    Pre p;
    p.o   = o;
    p.d   = d;
    p.inv = { 1.0f / Guard( d.x ), 1.0f / Guard( d.y ), 1.0f / Guard( d.z ) };
    const float ax = std::fabs( d.x ), ay = std::fabs( d.y ), az = std::fabs( d.z );
    int         kz = 0;
    if( ay > ax ) kz = 1;
    if( az > ( kz == 0 ? ax : ay ) ) kz = 2;
    int kx = ( kz + 1 ) % 3, ky = ( kx + 1 ) % 3;
    if( comp( d, kz ) < 0.0f ) std::swap( kx, ky );
    p.kx = kx; p.ky = ky; p.kz = kz;
    const float dz = comp( d, kz );
    p.Sx = comp( d, kx ) / dz;
    p.Sy = comp( d, ky ) / dz;
    p.Sz = 1.0f / dz;
    return p;
}

bool Slab( const Pre& r, const Node8& nd, int k, float tmin, float tmax, float* tnOut )
{
    const float tx0 = ( nd.lo_x[ k ] - r.o.x ) * r.inv.x, tx1 = ( nd.hi_x[ k ] - r.o.x ) * r.inv.x;
    const float ty0 = ( nd.lo_y[ k ] - r.o.y ) * r.inv.y, ty1 = ( nd.hi_y[ k ] - r.o.y ) * r.inv.y;
    const float tz0 = ( nd.lo_z[ k ] - r.o.z ) * r.inv.z, tz1 = ( nd.hi_z[ k ] - r.o.z ) * r.inv.z;
    const float tn = std::fmax( std::fmax( std::fmin( tx0, tx1 ), std::fmin( ty0, ty1 ) ), std::fmax( std::fmin( tz0, tz1 ), tmin ) );
    const float tf = std::fmin( std::fmin( std::fmax( tx0, tx1 ), std::fmax( ty0, ty1 ) ), std::fmax( tz0, tz1 ) ) * 1.0000004f;
    *tnOut = tn;
    return tn <= std::fmin( tf, tmax );
}

bool Tri( const Pre& r, F3 p0, F3 p1, F3 p2, float tmin, float tmax, bool cullFront, float* tO, float* uO, float* vO )
{
    const F3    A = sub( p0, r.o ), B = sub( p1, r.o ), C = sub( p2, r.o );
    const float Az = comp( A, r.kz ), Bz = comp( B, r.kz ), Cz = comp( C, r.kz );
    const float Ax = comp( A, r.kx ) - r.Sx * Az, Ay = comp( A, r.ky ) - r.Sy * Az;
    const float Bx = comp( B, r.kx ) - r.Sx * Bz, By = comp( B, r.ky ) - r.Sy * Bz;
    const float Cx = comp( C, r.kx ) - r.Sx * Cz, Cy = comp( C, r.ky ) - r.Sy * Cz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if( U == 0.0f || V == 0.0f || W == 0.0f )
    {
        U = float( double( Cx ) * double( By ) - double( Cy ) * double( Bx ) );
        V = float( double( Ax ) * double( Cy ) - double( Ay ) * double( Cx ) );
        W = float( double( Bx ) * double( Ay ) - double( By ) * double( Ax ) );
    }
    if( ( U < 0.0f || V < 0.0f || W < 0.0f ) && ( U > 0.0f || V > 0.0f || W > 0.0f ) ) return false;
    const float det = U + V + W;
    if( det == 0.0f ) return false;
    if( cullFront && det > 0.0f ) return false;
    const float T   = U * ( r.Sz * Az ) + V * ( r.Sz * Bz ) + W * ( r.Sz * Cz );
    const float rcp = 1.0f / det;
    const float t   = T * rcp;
    if( !( t >= tmin && t <= tmax ) ) return false;
    *tO = t;
    *uO = V * rcp;
    *vO = W * rcp;
    return true;
}
} // namespace

extern "C" void or_trace_rays( const OrBvh* b, const RgExtRay* rays, uint32_t count, RgExtHit* hits, uint32_t flags,
                               int threads, uint64_t* nodeVisits, uint64_t* triTests )
{
    or_trace_rays_filtered( b, rays, count, hits, flags, threads, nodeVisits, triTests, nullptr, nullptr );
}

// The same traversal with an any-hit filter: a candidate hit the filter
// rejects is ignored (ignoreIntersectionEXT of RtAlphaTest.rahit).
extern "C" void or_trace_rays_filtered( const OrBvh* b, const RgExtRay* rays, uint32_t count, RgExtHit* hits, uint32_t flags,
                                        int threads, uint64_t* nodeVisits, uint64_t* triTests, OrHitFilter accept, const void* ctx )
{
    const Node8* nodes     = reinterpret_cast< const Node8* >( b->nodes.data() );
    const float* tris      = b->sortedTris.data();
    const bool   anyHit    = ( flags & RG_EXT_TRACE_ANY_HIT ) != 0;
    const bool   cullFront = ( flags & RG_EXT_TRACE_CULL_FRONT_FACE ) != 0;
    std::atomic< uint64_t > nv{ 0 }, tt{ 0 };
    ParallelFor( count, threads, [ & ]( uint32_t i ) {
        const RgExtRay& ry = rays[ i ];
        const Pre r = MakePre( { ry.origin[ 0 ], ry.origin[ 1 ], ry.origin[ 2 ] }, { ry.direction[ 0 ], ry.direction[ 1 ], ry.direction[ 2 ] } );
        float    bestT = ry.tmax, bestU = 0, bestV = 0;
        uint32_t bestP = RG_EXT_MISS;
        uint64_t lnv = 0, ltt = 0;
        std::vector< uint32_t > stack;
        stack.reserve( 64 );
        stack.push_back( 0 );
        bool done = false;
        while( !stack.empty() && !done )
        {
            const uint32_t ref = stack.back();
            stack.pop_back();
            if( ( ref & LEAF_BIT ) == 0 )
            {
                lnv++;
                const Node8& nd = nodes[ ref ];
                float        dist[ 8 ];
                uint32_t     cr[ 8 ];
                int          h = 0;
                for( int k = 0; k < 8; k++ )
                {
                    float tn;
                    if( nd.child[ k ] != CHILD_EMPTY && Slab( r, nd, k, ry.tmin, bestT, &tn ) )
                    {
                        dist[ h ] = tn;
                        cr[ h ]   = nd.child[ k ];
                        h++;
                    }
                }
                // closest hit visits near to far: order by the 64-bit key
                // (distance as an order-preserving unsigned integer, +0 for
                // -0; then the child reference), push far first
                uint64_t key[ 8 ];
                for( int k = 0; k < h; k++ )
                {
                    const uint32_t u = bitcast< uint32_t >( dist[ k ] + 0.0f );
                    const uint32_t m = ( u & 0x80000000u ) ? ~u : ( u | 0x80000000u );
                    key[ k ]         = ( uint64_t( m ) << 32 ) | cr[ k ];
                }
                std::vector< int > ord( h );
                for( int k = 0; k < h; k++ ) ord[ k ] = k;
                // any hit: slot order, unsorted (the kernel's choice: the
                // interval does not shrink before the ray ends)
                if( !anyHit )
                {
                    std::sort( ord.begin(), ord.end(), [ & ]( int x, int y ) { return key[ x ] < key[ y ]; } );
                }
                for( int k = h - 1; k >= 0; k-- )
                {
                    stack.push_back( cr[ ord[ k ] ] );
                }
            }
            else
            {
                const uint32_t first = ( ref >> 3 ) & 0x0FFFFFFFu, cnt = ( ref & 7u ) + 1u;
                for( uint32_t k = 0; k < cnt; k++ )
                {
                    ltt++;
                    const float* q = tris + size_t( first + k ) * 12;
                    float        t, u, v;
                    if( Tri( r, { q[ 0 ], q[ 1 ], q[ 2 ] }, { q[ 4 ], q[ 5 ], q[ 6 ] }, { q[ 8 ], q[ 9 ], q[ 10 ] }, ry.tmin, bestT, cullFront, &t, &u, &v ) )
                    {
                        const uint32_t prim = bitcast< uint32_t >( q[ 3 ] );
                        if( accept && !accept( ctx, prim, u, v ) )
                        {
                            continue;
                        }
                        if( t < bestT || prim < bestP )
                        {
                            bestT = t; bestU = u; bestV = v; bestP = prim;
                            if( anyHit )
                            {
                                done = true;
                                break;
                            }
                        }
                    }
                }
            }
        }
        hits[ i ] = RgExtHit{ bestT, bestU, bestV, bestP };
        nv += lnv;
        tt += ltt;
    } );
    if( nodeVisits ) *nodeVisits = nv.load();
    if( triTests ) *triTests = tt.load();
}

// Moller & Trumbore, "Fast, Minimum Storage Ray/Triangle Intersection", 1997,
// in FP64 (independent of the watertight FP32 path).
extern "C" void or_trace_bruteforce_f64( const float* tris, uint32_t ntri, const RgExtRay* rays, uint32_t count,
                                         uint32_t flags, double* tOut, double* uOut, double* vOut, uint32_t* primOut,
                                         int threads )
{
    const bool cullFront = ( flags & RG_EXT_TRACE_CULL_FRONT_FACE ) != 0;
    ParallelFor( count, threads, [ & ]( uint32_t i ) {
        const RgExtRay& r = rays[ i ];
        const double    o[ 3 ] = { r.origin[ 0 ], r.origin[ 1 ], r.origin[ 2 ] };
        const double    d[ 3 ] = { r.direction[ 0 ], r.direction[ 1 ], r.direction[ 2 ] };
        double   bestT = r.tmax, bu = 0, bv = 0;
        uint32_t bp    = RG_EXT_MISS;
        for( uint32_t k = 0; k < ntri; k++ )
        {
            const float* t = tris + size_t( k ) * 9;
            double e1[ 3 ], e2[ 3 ], pv[ 3 ], tv[ 3 ], qv[ 3 ];
            for( int a = 0; a < 3; a++ )
            {
                e1[ a ] = double( t[ 3 + a ] ) - double( t[ a ] );
                e2[ a ] = double( t[ 6 + a ] ) - double( t[ a ] );
                tv[ a ] = o[ a ] - double( t[ a ] );
            }
            pv[ 0 ] = d[ 1 ] * e2[ 2 ] - d[ 2 ] * e2[ 1 ];
            pv[ 1 ] = d[ 2 ] * e2[ 0 ] - d[ 0 ] * e2[ 2 ];
            pv[ 2 ] = d[ 0 ] * e2[ 1 ] - d[ 1 ] * e2[ 0 ];
            const double det = e1[ 0 ] * pv[ 0 ] + e1[ 1 ] * pv[ 1 ] + e1[ 2 ] * pv[ 2 ];
            if( det == 0.0 ) continue;
            // front-facing (counter-clockwise from the origin) <=> dot(n, d) < 0
            // <=> det = dot(d, e2 x e1)... = -dot(n, d) > 0
            if( cullFront && det > 0.0 ) continue;
            const double inv = 1.0 / det;
            const double u   = ( tv[ 0 ] * pv[ 0 ] + tv[ 1 ] * pv[ 1 ] + tv[ 2 ] * pv[ 2 ] ) * inv;
            if( u < 0.0 || u > 1.0 ) continue;
            qv[ 0 ] = tv[ 1 ] * e1[ 2 ] - tv[ 2 ] * e1[ 1 ];
            qv[ 1 ] = tv[ 2 ] * e1[ 0 ] - tv[ 0 ] * e1[ 2 ];
            qv[ 2 ] = tv[ 0 ] * e1[ 1 ] - tv[ 1 ] * e1[ 0 ];
            const double v = ( d[ 0 ] * qv[ 0 ] + d[ 1 ] * qv[ 1 ] + d[ 2 ] * qv[ 2 ] ) * inv;
            if( v < 0.0 || u + v > 1.0 ) continue;
            const double tt = ( e2[ 0 ] * qv[ 0 ] + e2[ 1 ] * qv[ 1 ] + e2[ 2 ] * qv[ 2 ] ) * inv;
            if( tt < double( r.tmin ) || tt > bestT ) continue;
            if( tt < bestT || k < bp )
            {
                bestT = tt; bu = u; bv = v; bp = k;
            }
        }
        tOut[ i ] = bestT; uOut[ i ] = bu; vOut[ i ] = bv; primOut[ i ] = bp;
    } );
}

// =====================================================================
// Eye adaptation: CmLuminanceHistogram.comp:41-86, CmLuminanceAvg.comp:39-90.
// =====================================================================
extern "C" void or_lum_histogram( const float* lum, uint32_t count, float minLog, float maxLog, uint32_t* hist )
{
    const float inv = 1.0f / ( maxLog - minLog );
    for( uint32_t i = 0; i < count; i++ )
    {
        const float l = lum[ i ];
        uint32_t    bin;
        if( !( l >= 0.001f ) )
        {
            bin = 0; // LUMINANCE_EPS (CmLuminanceHistogram.comp:22); NaN too
        }
        else if( l > 3.0e38f )
        {
            bin = 255;
        }
        else
        {
            float x = ( or_log2f( l ) - minLog ) * inv;
            x       = std::fmin( std::fmax( x, 0.0f ), 1.0f );
            bin     = uint32_t( x * 254.0f + 1.0f );
        }
        hist[ bin ]++;
    }
}

extern "C" float or_lum_average( uint32_t* hist, uint32_t pixelCount, float minLog, float maxLog, float timeDelta, int stop, float lastAvg )
{
    uint32_t weighted = 0;
    for( uint32_t i = 0; i < 256; i++ )
    {
        weighted += hist[ i ] * i;
    }
    const uint32_t black = hist[ 0 ];
    for( uint32_t i = 0; i < 256; i++ )
    {
        hist[ i ] = 0;
    }
    if( stop )
    {
        return lastAvg;
    }
    const float range   = maxLog - minLog;
    const float w       = ( float( weighted ) / std::fmax( float( pixelCount ) - float( black ), 1.0f ) ) - 1.0f;
    const float avg     = or_exp2f( w / 254.0f * range + minLog );
    const float t       = 1.0f - or_exp2f( -timeDelta * 1.1f * 1.44269504f );
    const float adapted = lastAvg + ( avg - lastAvg ) * t;
    return std::fmin( adapted, or_exp2f( maxLog ) );
}
