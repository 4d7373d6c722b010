// Internal declarations shared by the HIP translation units of librtgl1_amd.
//
// Error model: internal code throws rg::Error (a result code and a message);
// every extern "C" entry point catches it and returns the code, as the
// reference's Call() wrapper converts RgException to RgResult
// (Source/RTGL1.cpp:146-194).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "RTGL1/RTGL1.h"
#include "RTGL1/RTGL1_ext.h"

namespace rg
{

struct Error : std::runtime_error
{
    RgResult code;
    Error( RgResult c, const std::string& msg ) : std::runtime_error( msg ), code( c ) {}
};

void SetLastError( const char* msg );

#define RG_HIP_CHECK( expr )                                                              \
    do                                                                                    \
    {                                                                                     \
        hipError_t rg_e_ = ( expr );                                                      \
        if( rg_e_ != hipSuccess )                                                         \
        {                                                                                 \
            throw ::rg::Error( RG_RESULT_GRAPHICS_API_ERROR,                              \
                               std::string( #expr ) + ": " + hipGetErrorString( rg_e_ ) ); \
        }                                                                                 \
    } while( 0 )

#define RG_REQUIRE( cond, code, msg )                   \
    do                                                  \
    {                                                   \
        if( !( cond ) )                                 \
        {                                               \
            throw ::rg::Error( ( code ), ( msg ) );     \
        }                                               \
    } while( 0 )

// Runs f() and converts exceptions into RgResult at the C boundary.
template< typename F >
RgResult Guard( F&& f )
{
    try
    {
        f();
        SetLastError( "" );
        return RG_RESULT_SUCCESS;
    }
    catch( const Error& e )
    {
        SetLastError( e.what() );
        return e.code;
    }
    catch( const std::bad_alloc& )
    {
        SetLastError( "host allocation failed" );
        return RG_RESULT_INTERNAL_ERROR;
    }
    catch( const std::exception& e )
    {
        SetLastError( e.what() );
        return RG_RESULT_INTERNAL_ERROR;
    }
}

inline hipStream_t AsStream( void* s )
{
    return reinterpret_cast< hipStream_t >( s );
}

inline uint32_t DivUp( uint64_t a, uint64_t b )
{
    return uint32_t( ( a + b - 1 ) / b );
}

// Checks for a kernel launch error right after a <<<>>> launch.
inline void CheckLaunch( const char* what )
{
    hipError_t e = hipGetLastError();
    if( e != hipSuccess )
    {
        throw Error( RG_RESULT_GRAPHICS_API_ERROR, std::string( what ) + ": " + hipGetErrorString( e ) );
    }
}

// A growable device buffer.
struct DeviceBuffer
{
    void*  ptr   = nullptr;
    size_t bytes = 0;

    DeviceBuffer() = default;
    DeviceBuffer( const DeviceBuffer& ) = delete;
    DeviceBuffer& operator=( const DeviceBuffer& ) = delete;
    ~DeviceBuffer() { Release(); }

    void Release()
    {
        if( ptr )
        {
            ( void )hipFree( ptr );
            ptr   = nullptr;
            bytes = 0;
        }
    }

    // Ensures at least `need` bytes; contents are not preserved on growth.
    void Reserve( size_t need )
    {
        if( need <= bytes )
        {
            return;
        }
        Release();
        size_t n = need < 256 ? 256 : need;
        RG_HIP_CHECK( hipMalloc( &ptr, n ) );
        bytes = n;
    }

    template< typename T >
    T* as() const
    {
        return reinterpret_cast< T* >( ptr );
    }
};

} // namespace rg
