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

#include "arwen_dycore.h"

typedef arw_domain Dom;

// Float32 scalars derived from the double-precision configuration exactly
// as the reference's host code derives them before a launch
// (rdx = float32(1/dx), dtau = float32(dt/ns), dampmag =
// float32(dtau_double * dampcoef), ...).
struct Sc {
    float rdx, rdy;          // float32(1/dx)
    float hrdx, hrdy;        // float32(0.5/dx)
    float qrdx, qrdy;        // float32(0.25/dx)
    float dxi2, dyi2;        // float32(1/dx**2)
    float kh, kv;
    float dtau, epssm, smdiv;
    float dampmag, zdamp;
    int top_lid, diff, damp, first, use_pold;
};

static inline Sc make_sc(const arw_params* p, double dtau, int first)
{
    Sc s;
    double rdx = 1.0 / p->dx, rdy = 1.0 / p->dy;
    s.rdx = (float)rdx;
    s.rdy = (float)rdy;
    s.hrdx = (float)(0.5 * rdx);
    s.hrdy = (float)(0.5 * rdy);
    s.qrdx = (float)(0.25 * rdx);
    s.qrdy = (float)(0.25 * rdy);
    s.dxi2 = (float)(1.0 / (p->dx * p->dx));
    s.dyi2 = (float)(1.0 / (p->dy * p->dy));
    s.kh = (float)p->khdif;
    s.kv = (float)p->kvdif;
    s.dtau = (float)dtau;
    s.epssm = (float)p->epssm;
    s.smdiv = first ? 0.0f : (float)p->smdiv;
    s.first = first != 0;
    s.use_pold = 1;
    s.damp = p->damp_opt == 3;
    s.dampmag = s.damp ? (float)(dtau * p->dampcoef) : 0.0f;
    s.zdamp = (float)p->zdamp;
    s.top_lid = p->top_lid != 0;
    s.diff = (p->khdif > 0.0 || p->kvdif > 0.0);
    return s;
}

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

// ---------------------------------------------------------------------------
// Column-march addressing.
//
// A column kernel touches the same few columns at every level, so each load
// address splits into a per-thread column byte offset (fixed for the whole
// march, one 32-bit VGPR) and a per-level offset that is the same for every
// lane.  gld/gst form base + level in scalar registers and add the 32-bit
// column offset in the load itself (global_load_dword v, voff, s[base:+1]),
// so the inner loops carry no 64-bit per-lane address arithmetic.  The
// compiler only selects that form when it sees the 32-bit offset being
// zero-extended in the same block as the load; colb_opaque() at the top of
// each loop iteration keeps the offsets from being widened once, outside
// the loop.  Byte offsets are unsigned 32-bit: the host keeps every array
// below 4 GiB.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gld(const float* base, unsigned lev,
                                     unsigned colb)
{
    const char* b = reinterpret_cast<const char*>(base + lev);
    return *reinterpret_cast<const float*>(b + (uint64_t)colb);
}

__device__ __forceinline__ void gst(float* base, unsigned lev, unsigned colb,
                                    float x)
{
    char* b = reinterpret_cast<char*>(base + lev);
    *reinterpret_cast<float*>(b + (uint64_t)colb) = x;
}

__device__ __forceinline__ void colb_opaque(unsigned& a)
{
    asm volatile("" : "+v"(a));
}

// 1-D level coefficients (c1h, rdnw, fnm, ... and the 1-D base profiles)
// are never written by any kernel.  Reading them through the constant
// address space lets a uniform index become a scalar load: it is tracked
// by lgkmcnt instead of the in-order vector-memory counter, so waiting for
// a coefficient does not also wait for the next level's prefetched loads.
typedef const __attribute__((address_space(4))) float kfloat;

__device__ __forceinline__ kfloat* kview(const float* p)
{
    return (kfloat*)p;
}

// ---------------------------------------------------------------------------
// XCD-aware workgroup numbering.  MI355X hands workgroups to its 8 XCDs
// round-robin, and each XCD has its own L2.  Column kernels read rows j-1
// and j+1 (or j-3..j+3) that other workgroups of the same launch also
// read; renumbering so that XCD x runs the x-th contiguous chunk of the
// domain keeps those neighbouring rows in one L2.  The host pads the grid
// to a multiple of ARW_NXCD; padded workgroups fall outside the domain and
// exit at the bounds check.
// ---------------------------------------------------------------------------
#define ARW_NXCD 8

__device__ __forceinline__ int xcd_block()
{
    const int n = (int)gridDim.x, b = (int)blockIdx.x;
    return (b % ARW_NXCD) * (n / ARW_NXCD) + b / ARW_NXCD;
}

static inline unsigned xcd_grid(unsigned nblocks)
{
    return (nblocks + ARW_NXCD - 1) / ARW_NXCD * ARW_NXCD;
}

// Pin a value as computed unconditionally.  A select between a cheap and
// an expensive operand (a division) is otherwise turned back into a
// branch by the code generator, splitting a level loop into blocks.
__device__ __forceinline__ void keep(float& x)
{
    asm volatile("" : "+v"(x));
}
