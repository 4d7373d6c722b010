// Ray/box and watertight ray/triangle tests shared by every traversal kernel.
#pragma once

#include "rg_device_math.h"

namespace rg
{

// Per-ray constants of the watertight test (Woop, Benthin, Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013): the axis of largest |d| is z,
// the other two are sheared so that the ray becomes (0, 0, 1).
struct RayPre
{
    v3    o;
    v3    d;
    v3    invd;
    int   kx, ky, kz;
    float Sx, Sy, Sz;
};

__host__ __device__ __forceinline__ float GuardDir( float d )
{
    // Keeps 1/d finite so that (lo - o) * (1/d) is never 0 * inf.
    return fabsf( d ) >= 1e-20f ? d : ( d < 0.0f ? -1e-20f : 1e-20f );
}

__host__ __device__ __forceinline__ RayPre MakeRayPre( v3 o, v3 d )
{
    RayPre r;
    r.o    = o;
    r.d    = d;
    r.invd = mk3( 1.0f / GuardDir( d.x ), 1.0f / GuardDir( d.y ), 1.0f / GuardDir( d.z ) );

    const float ax = fabsf( d.x ), ay = fabsf( d.y ), az = fabsf( d.z );
    int         kz = 0;
    if( ay > ax )
    {
        kz = 1;
    }
    if( az > ( kz == 0 ? ax : ay ) )
    {
        kz = 2;
    }
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if( get( d, kz ) < 0.0f )
    {
        int t = kx;
        kx    = ky;
        ky    = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    const float dz = get( d, kz );
    r.Sx           = get( d, kx ) / dz;
    r.Sy           = get( d, ky ) / dz;
    r.Sz           = 1.0f / dz;
    return r;
}

// Robust slab test (Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013): the far
// distance is scaled by 1 + 2*gamma(3) so rounding never culls a box that the
// exact ray touches.
__host__ __device__ __forceinline__ bool SlabTest( const RayPre& r,
                                                   float lx, float ly, float lz,
                                                   float hx, float hy, float hz,
                                                   float tmin, float tmax, float* tnearOut )
{
    const float tx0 = ( lx - r.o.x ) * r.invd.x, tx1 = ( hx - r.o.x ) * r.invd.x;
    const float ty0 = ( ly - r.o.y ) * r.invd.y, ty1 = ( hy - r.o.y ) * r.invd.y;
    const float tz0 = ( lz - r.o.z ) * r.invd.z, tz1 = ( hz - r.o.z ) * r.invd.z;
    const float tn  = fmaxf( fmaxf( fminf( tx0, tx1 ), fminf( ty0, ty1 ) ), fmaxf( fminf( tz0, tz1 ), tmin ) );
    const float tf  = fminf( fminf( fmaxf( tx0, tx1 ), fmaxf( ty0, ty1 ) ), fmaxf( tz0, tz1 ) ) * 1.0000004f;
    *tnearOut       = tn;
    return tn <= fminf( tf, tmax );
}

// SlabTest with the planes already ordered for the ray: n* is the box side
// the ray enters through on each axis (lo where 1/d > 0, hi where 1/d < 0),
// f* the side it leaves through.  (b - o) * (1/d) rounds monotonically in
// b, so for lo <= hi the entering side's distance is the min of the two and
// the leaving side's the max: the same tn and tf as SlabTest, bit for bit
// (up to the sign of a zero tn, which no comparison or sort key sees),
// without the six min / max per box.
__host__ __device__ __forceinline__ bool SlabTestOrdered( const RayPre& r,
                                                          float nx, float ny, float nz,
                                                          float fx, float fy, float fz,
                                                          float tmin, float tmax, float* tnearOut )
{
    const float tx0 = ( nx - r.o.x ) * r.invd.x, tx1 = ( fx - r.o.x ) * r.invd.x;
    const float ty0 = ( ny - r.o.y ) * r.invd.y, ty1 = ( fy - r.o.y ) * r.invd.y;
    const float tz0 = ( nz - r.o.z ) * r.invd.z, tz1 = ( fz - r.o.z ) * r.invd.z;
    const float tn  = fmaxf( fmaxf( tx0, ty0 ), fmaxf( tz0, tmin ) );
    const float tf  = fminf( fminf( tx1, ty1 ), tz1 ) * 1.0000004f;
    *tnearOut       = tn;
    return tn <= fminf( tf, tmax );
}

// Watertight intersection in [tmin, tmax] (both inclusive).  u, v are the
// barycentric weights of p1 and p2.  cullFront rejects triangles whose
// vertices appear counter-clockwise from the ray origin (dot(n, d) < 0 for
// n = (p1 - p0) x (p2 - p0)), i.e. gl_RayFlagsCullFrontFacingTrianglesEXT
// with counter-clockwise front faces.
__host__ __device__ __forceinline__ bool IntersectTri( const RayPre& r, v3 p0, v3 p1, v3 p2,
                                                       float tmin, float tmax, bool cullFront,
                                                       float* tOut, float* uOut, float* vOut )
{
    const v3    A  = p0 - r.o, B = p1 - r.o, C = p2 - r.o;
    const float Az = get( A, r.kz ), Bz = get( B, r.kz ), Cz = get( C, r.kz );
    const float Ax = get( A, r.kx ) - r.Sx * Az, Ay = get( A, r.ky ) - r.Sy * Az;
    const float Bx = get( B, r.kx ) - r.Sx * Bz, By = get( B, r.ky ) - r.Sy * Bz;
    const float Cx = get( C, r.kx ) - r.Sx * Cz, Cy = get( C, r.ky ) - r.Sy * Cz;

    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if( U == 0.0f || V == 0.0f || W == 0.0f )
    {
        // Edge case: recompute the 2x2 determinants in FP64 (the paper's fallback).
        U = float( double( Cx ) * double( By ) - double( Cy ) * double( Bx ) );
        V = float( double( Ax ) * double( Cy ) - double( Ay ) * double( Cx ) );
        W = float( double( Bx ) * double( Ay ) - double( By ) * double( Ax ) );
    }
    if( ( U < 0.0f || V < 0.0f || W < 0.0f ) && ( U > 0.0f || V > 0.0f || W > 0.0f ) )
    {
        return false;
    }
    const float det = U + V + W;
    if( det == 0.0f )
    {
        return false;
    }
    if( cullFront && det > 0.0f )
    {
        return false;
    }
    const float T   = U * ( r.Sz * Az ) + V * ( r.Sz * Bz ) + W * ( r.Sz * Cz );
    const float rcp = 1.0f / det;
    const float t   = T * rcp;
    if( !( t >= tmin && t <= tmax ) )
    {
        return false;
    }
    *tOut = t;
    *uOut = V * rcp;
    *vOut = W * rcp;
    return true;
}

} // namespace rg
