// Shared device helpers for the gfx950 dycore kernels.
//
// Arithmetic contract: the library is compiled with -ffp-contract=off and
// every expression below is written in the evaluation order of the
// float32 NumPy restatement of the reference step, so the device result
// rounds exactly like that restatement.  Powers and sines are evaluated
// in double precision and rounded once to float.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define G_F 9.81f
#define RD_F 287.0f
#define P0_F 100000.0f
#define GAMMA_F 1.4f
#define RVOVRD_F 1.6083624362945557f

// Positive modulo for small offsets and any n >= 1.
__device__ __forceinline__ int pmod(int i, int n)
{
    int r = i % n;
    return r < 0 ? r + n : r;
}

// Flat indices.  Mass/w fields: (k, j, i) with row width nx; u: nx+1;
// v: ny+1 rows.  Extents are checked on the host to fit in int32.
__device__ __forceinline__ int iM(int k, int j, int i, int ny, int nx)
{
    return (k * ny + j) * nx + i;
}
__device__ __forceinline__ int iU(int k, int j, int i, int ny, int nx)
{
    return (k * ny + j) * (nx + 1) + i;
}
__device__ __forceinline__ int iV(int k, int j, int i, int ny, int nx)
{
    return (k * (ny + 1) + j) * nx + i;
}

// WRF flux5 (upwind-biased 5th order); face between qm1 and q0.
__device__ __forceinline__ float flux5(float qm3, float qm2, float qm1,
                                       float q0, float qp1, float qp2,
                                       float vel)
{
    float a = 37.0f * (q0 + qm1) - 8.0f * (qp1 + qm2) + (qp2 + qm3);
    float b = 10.0f * (q0 - qm1) - 5.0f * (qp1 - qm2) + (qp2 - qm3);
    return (vel * a - fabsf(vel) * b) / 60.0f;
}

// Omega-signed vertical flux3 (eta decreases upward).
__device__ __forceinline__ float flux3(float qm2, float qm1, float q0,
                                       float qp1, float vel)
{
    float a = 7.0f * (q0 + qm1) - (qp1 + qm2);
    float b = 3.0f * (q0 - qm1) - (qp1 - qm2);
    return (vel * a + fabsf(vel) * b) / 12.0f;
}

// x**GAMMA evaluated in double and rounded once.
__device__ __forceinline__ float pow_gamma(float x)
{
    return (float)pow((double)x, (double)GAMMA_F);
}

// Base-profile value: 1-D column (base3d = 0) or per-point 3-D field.
__device__ __forceinline__ float basev(const float* __restrict__ b, int k,
                                       int c, int st, int base3d)
{
    return base3d ? b[k * st + c] : b[k];
}
