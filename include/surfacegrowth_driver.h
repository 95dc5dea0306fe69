/*
 * surfacegrowth_driver.h - what the reference's host driver needs in place of its CUDA header.
 *
 * The reference host driver (SurfaceGrowthConsole/SurfaceGrowth.cpp with SurfaceGrowthProto.h)
 * reaches the engine through SurfaceGrowth.h, which pulls in the CUDA runtime, the Windows
 * <tchar.h>/<conio.h> and ComputeDefs.h (itself on CUDA's vector_types.h).  Pointing
 * SurfaceGrowthProto.h:11 at this header instead gives the driver:
 *   - SimParams, TBuf, MATERIAL, Prop, real and the four wrappers, from surfacegrowth.h;
 *   - the names float3/float4/int3/uint3/uint for the layout-identical PODs of surfacegrowth.h,
 *     and the host vectors VecR/VecI with NDIM (ComputeDefs.h:26-48);
 *   - the non-Unicode Windows names the driver uses (TCHAR, _TCHAR, TEXT, BOOL, lstrcpy, lstrcat;
 *     SurfaceGrowth.h:40-62);
 *   - BLOCK_SIZE and the LCG constants (SurfaceGrowth.h:35-41);
 *   - the Rapaport vector and property macros the driver calls (VZero, VCopy, VScale, VSCopy,
 *     VVAdd, VVSAdd, VDiv, Sqr, AllocMem, PropZero; ComputeDefs.h:53-180), same argument order.
 * The driver's three device queries (cudaGetDeviceCount, cudaGetDeviceProperties and the
 * maxGridSize test, SurfaceGrowth.cpp:12, 109-122, 317) are edits to the driver: they become
 * sgDeviceCount() and sgMaxCellGrid(), declared in surfacegrowth.h.  No CUDA name is defined here.
 */
#ifndef SURFACEGROWTH_DRIVER_H
#define SURFACEGROWTH_DRIVER_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "surfacegrowth.h"

typedef sg_float3 float3;
typedef sg_float4 float4;
typedef sg_int3 int3;
typedef sg_uint3 uint3;
typedef unsigned int uint;

/* Host vector types of the driver's globals (ComputeDefs.h:26-48). */
#ifndef NDIM
#define NDIM 3
#endif
typedef struct VecR { real x, y, z; } VecR;
typedef struct VecI { int x, y, z; } VecI;

/* Non-Unicode Windows names (SurfaceGrowth.h:40-62). */
typedef char TCHAR;
typedef char _TCHAR;
#ifndef TEXT
#define TEXT(s) s
#endif
#define BOOL int
#define FALSE 0
#define TRUE 1
#define INT int
#define lstrcpy strcpy
#define lstrcat strcat

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Atoms per cell of the reference's cell array (SurfaceGrowth.h:35). */
#define BLOCK_SIZE SG_BLOCK_SIZE

/* Rapaport's LCG constants (SurfaceGrowth.h:37-41), used by the driver's RandR. */
#define IADD 453806245
#define IMUL 314159269
#define MASK 2147483647
#define SCALE 0.4656612873e-9

/* Vector and property macros, component-wise over .x .y .z (Rapaport's naming). */
#define Sqr(a) ((a) * (a))
#define Cube(a) ((a) * (a) * (a))
#define Max(a, b) ((a) > (b) ? (a) : (b))
#define Min(a, b) ((a) < (b) ? (a) : (b))
#define AllocMem(p, n, T) p = (T *)malloc((n) * sizeof(T))
#define VSet(v, a, b, c) ((v).x = (a), (v).y = (b), (v).z = (c))
#define VZero(v) VSet(v, 0, 0, 0)
#define VCopy(d, s) VSet(d, (s).x, (s).y, (s).z)
#define VScale(v, k) ((v).x *= (k), (v).y *= (k), (v).z *= (k))
#define VSCopy(d, k, s) VSet(d, (k) * (s).x, (k) * (s).y, (k) * (s).z)
#define VAdd(d, a, b) VSet(d, (a).x + (b).x, (a).y + (b).y, (a).z + (b).z)
#define VSub(d, a, b) VSet(d, (a).x - (b).x, (a).y - (b).y, (a).z - (b).z)
#define VDiv(d, a, b) VSet(d, (a).x / (b).x, (a).y / (b).y, (a).z / (b).z)
#define VSAdd(d, a, k, b) VSet(d, (a).x + (k) * (b).x, (a).y + (k) * (b).y, (a).z + (k) * (b).z)
#define VVAdd(d, b) VAdd(d, d, b)
#define VVSub(d, b) VSub(d, d, b)
#define VVSAdd(d, k, b) VSAdd(d, d, k, b)
#define VDot(a, b) ((a).x * (b).x + (a).y * (b).y + (a).z * (b).z)
#define VLenSq(v) VDot(v, v)
#define VProd(v) ((v).x * (v).y * (v).z)
#define PropZero(p) ((p).sum = (p).sum2 = 0.)
#define PropAccum(p) ((p).sum += (p).val, (p).sum2 += Sqr((p).val))
#define PropAvg(p, n) ((p).sum /= (n), (p).sum2 = sqrt(Max((p).sum2 / (n) - Sqr((p).sum), 0.)))

#endif /* SURFACEGROWTH_DRIVER_H */
