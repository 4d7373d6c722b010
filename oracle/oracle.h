/*
 * CPU ORACLE - TEST INFRASTRUCTURE ONLY.
 *
 * A scalar CPU restatement of the RTGL1 hot path, used by tests/ as the
 * parity checker, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg.  The shipped library (raytracedgl1_amd/librtgl1_amd.so) never links,
 * loads or calls anything here.
 *
 * Each function cites the reference file it restates.  Where the reference
 * delegates to the Vulkan driver (acceleration-structure build and
 * traversal), the oracle restates the published algorithm the GPU build
 * implements and is pinned by invariants and by an independent FP64
 * brute-force intersector instead.
 */
#ifndef RG_ORACLE_H_
#define RG_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#include "RTGL1/RTGL1_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- portable math (bit-identical to the kernels' rgLog2 / rgExp2) ---- */
float or_log2f( float x );
float or_exp2f( float x );

/* ---- codecs and RNG (Source/Shaders/Utils.h, Random.h) ---- */
uint32_t or_wellons_low_bias32( uint32_t x );
void     or_murmur_hash33( uint32_t x, uint32_t y, uint32_t z, uint32_t out[ 3 ] );
uint32_t or_random_seed( int px, int py, uint32_t frameIndex );
uint32_t or_encode_normal( const double n[ 3 ] );
void     or_decode_normal( uint32_t packed, double n[ 3 ] );
uint32_t or_encode_e5b9g9r9( const double rgb[ 3 ] );
void     or_decode_e5b9g9r9( uint32_t packed, double rgb[ 3 ] );
void     or_checkerboard_from_regular( int x, int y, int width, int* cx, int* cy );
void     or_checkerboard_to_regular( int cx, int cy, int width, int* x, int* y );

/* ---- LBVH (no reference code: restates the published algorithms) ---- */
typedef struct OrBvh OrBvh;
OrBvh*   or_bvh_build( const float* tris, uint32_t n, uint32_t leafMax );
void     or_bvh_free( OrBvh* b );
uint32_t or_bvh_node_count( const OrBvh* b );
void     or_bvh_scene_bounds( const OrBvh* b, double outMin[ 3 ], double outMax[ 3 ] );
void     or_bvh_copy_codes( const OrBvh* b, uint64_t* out );          /* input order      */
void     or_bvh_copy_sorted_indices( const OrBvh* b, uint32_t* out ); /* n                */
void     or_bvh_copy_binary( const OrBvh* b, uint32_t* children, float* bounds ); /* 2(n-1), 6(2n-1) */
void     or_bvh_copy_nodes( const OrBvh* b, void* out );               /* nodeCount * 256  */

/* ---- traversal ---- */
/* Same watertight test and closest-hit rule as the GPU kernel, multi-threaded
 * over rays.  Optional statistics: total node visits and triangle tests. */
void or_trace_rays( const OrBvh* b, const RgExtRay* rays, uint32_t count, RgExtHit* hits, uint32_t flags,
                    int threads, uint64_t* nodeVisits, uint64_t* triTests );
/* With an any-hit filter: candidate hits for which accept returns 0 are
 * ignored (the alpha test). */
typedef int ( *OrHitFilter )( const void* ctx, uint32_t prim, float u, float v );
void or_trace_rays_filtered( const OrBvh* b, const RgExtRay* rays, uint32_t count, RgExtHit* hits, uint32_t flags,
                             int threads, uint64_t* nodeVisits, uint64_t* triTests, OrHitFilter accept, const void* ctx );
/* Independent FP64 Moller-Trumbore over every triangle (closest hit, ties to
 * the smaller id).  t, u, v in double. */
void or_trace_bruteforce_f64( const float* tris, uint32_t ntri, const RgExtRay* rays, uint32_t count,
                              uint32_t flags, double* tOut, double* uOut, double* vOut, uint32_t* primOut,
                              int threads );

/* ---- eye adaptation (CmLuminanceHistogram.comp, CmLuminanceAvg.comp) ---- */
void or_lum_histogram( const float* lum, uint32_t count, float minLog, float maxLog, uint32_t* hist );
/* Returns the new average luminance; clears hist. */
float or_lum_average( uint32_t* hist, uint32_t pixelCount, float minLog, float maxLog, float timeDelta,
                      int stop, float lastAvg );

#ifdef __cplusplus
}
#endif

#endif /* RG_ORACLE_H_ */
