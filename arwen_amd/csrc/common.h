// Shared device helpers for the gfx950 dycore kernels.
//
// Arithmetic contract: the library is compiled with -ffp-contract=off and
// every expression below is written in the evaluation order of the NumPy
// restatement of the reference step, so the device result rounds exactly
// like that restatement in the same precision.  In the float32 build,
// powers and sines are evaluated in double precision and rounded once.
//
// Precision: every kernel source is compiled twice, with ARW_F64 = 0
// (real = float, entry points arw_*) and ARW_F64 = 1 (real = double,
// entry points arw64_*): the FP64 build is the same discretisation in
// double precision, the reference the float32 step's error is measured
// against.  R(x) spells a literal in the build's precision (R(0.5) is
// 0.5f in the float32 build, so its instructions are unchanged).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "arwen_dycore.h"

#if ARW_F64
typedef double real;
typedef arw64_domain Dom;
typedef arw64_tracer_ckpt Ckpt;
#define R(x) x
#define ARW_NS prec64
#define ARW_API(name) arw64_##name
#else
typedef float real;
typedef arw_domain Dom;
typedef arw_tracer_ckpt Ckpt;
#define R(x) x##f
#define ARW_NS prec32
#define ARW_API(name) arw_##name
#endif
// bytes per element: the unit of the column-march byte offsets
#define RB ((unsigned)sizeof(real))

// Scalars derived from the double-precision configuration exactly as the
// reference's host code derives them before a launch (rdx = real(1/dx),
// dtau = real(dt/ns), dampmag = real(dtau_double * dampcoef), ...).
struct Sc {
    real rdx, rdy;          // float32(1/dx)
    real hrdx, hrdy;        // float32(0.5/dx)
    real qrdx, qrdy;        // float32(0.25/dx)
    real dxi2, dyi2;        // float32(1/dx**2)
    real kh, kv;
    real dtau, epssm, smdiv;
    real dampmag, zdamp;
    int top_lid, diff, damp, first, use_pold, prec;
    // k_column: the stage's Thomas factors gamma are already stored (they
    // depend on the stage reference only, so later substeps of a stage
    // skip storing them); park = levels at the top of each column whose
    // gamma and eliminated w'' stay in LDS between the two sweeps
    int gam_ready, park;
    // WRF sumflux fused into the substep pass: 0 off, 1 the stage's first
    // substep (the sums start from zero), 2 later substeps (accumulate);
    // k_uv adds u'', v'' into ru_m, rv_m and k_column Omega'' into ww_m
    int sum;
    // rhs_ph's horizontal geopotential advection: the 5th-order form
    int h5;
    // moist_cq: the pressure terms take d.cqu / d.cqv / d.cqw
    int cq;
    // held mixing tendencies (d.d6_*) join the stage tendencies as they are
    // stored (arw_step only)
    int held;
};

static inline Sc make_sc(const arw_params* p, double dtau, int first)
{
    Sc s;
    double rdx = 1.0 / p->dx, rdy = 1.0 / p->dy;
    s.rdx = (real)rdx;
    s.rdy = (real)rdy;
    s.hrdx = (real)(0.5 * rdx);
    s.hrdy = (real)(0.5 * rdy);
    s.qrdx = (real)(0.25 * rdx);
    s.qrdy = (real)(0.25 * rdy);
    s.dxi2 = (real)(1.0 / (p->dx * p->dx));
    s.dyi2 = (real)(1.0 / (p->dy * p->dy));
    s.kh = (real)p->khdif;
    s.kv = (real)p->kvdif;
    s.dtau = (real)dtau;
    s.epssm = (real)p->epssm;
    s.smdiv = first ? R(0.0) : (real)p->smdiv;
    s.first = first != 0;
    s.use_pold = 1;
    s.damp = p->damp_opt == 3;
    s.dampmag = s.damp ? (real)(dtau * p->dampcoef) : R(0.0);
    s.zdamp = (real)p->zdamp;
    s.top_lid = p->top_lid != 0;
    s.diff = (p->khdif > 0.0 || p->kvdif > 0.0);
    s.prec = (!ARW_F64) && p->precision_form != 0;
    s.gam_ready = 0;
    s.park = 0;
    s.sum = 0;
    s.h5 = p->h_sca_adv_order == 5;
    s.cq = p->moist_cq != 0;
    s.held = 0;
    return s;
}

#define G_F R(9.81)
#define RD_F R(287.0)
#define P0_F R(100000.0)
#define GAMMA_F R(1.4)
#define RVOVRD_F R(1.6083624362945557)

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

// x / c for the flux denominators c = 60, 12, correctly rounded.  In the
// FP32 build the quotient is the double product x * (1/c) rounded once to
// float (three instructions instead of a full division): the exact
// quotient's binary tail beyond float precision is a repeating block
// (1/15 = 0.(0001), 1/3 = 0.(01) in binary), so unless the quotient is
// exact it never lies within the double product's error of a float
// rounding midpoint, and the result is the correctly rounded float
// quotient.  Checked against true division for every float input.
__device__ __forceinline__ real div_const(real x, double inv, real c)
{
#if ARW_F64
    (void)inv;
    return x / c;
#else
    (void)c;
    return (real)((double)x * inv);
#endif
}

// WRF flux5 (upwind-biased 5th order); face between qm1 and q0.
__device__ __forceinline__ real flux5(real qm3, real qm2, real qm1,
                                       real q0, real qp1, real qp2,
                                       real vel)
{
    real a = R(37.0) * (q0 + qm1) - R(8.0) * (qp1 + qm2) + (qp2 + qm3);
    real b = R(10.0) * (q0 - qm1) - R(5.0) * (qp1 - qm2) + (qp2 - qm3);
    return div_const(vel * a - fabs(vel) * b, 1.0 / 60.0, R(60.0));
}

// Omega-signed vertical flux3 (eta decreases upward).
__device__ __forceinline__ real flux3(real qm2, real qm1, real q0,
                                       real qp1, real vel)
{
    real a = R(7.0) * (q0 + qm1) - (qp1 + qm2);
    real b = R(3.0) * (q0 - qm1) - (qp1 - qm2);
    return div_const(vel * a + fabs(vel) * b, 1.0 / 12.0, R(12.0));
}

// x**GAMMA evaluated in double and rounded once.
__device__ __forceinline__ real pow_gamma(real x)
{
    return (real)pow((double)x, (double)GAMMA_F);
}

// The EOS of one cell in perturbation-from-base form (the precision form).
// The base state is discretely hydrostatic (alb = -rdnw dphb / (c1h mub +
// c2h)) and in its own EOS (pb = P0 (RD thb / (P0 alb))^GAMMA), so with
// mh = c1h (mub + mu') + c2h
//   al = -(alb c1h mu' + rdnw dphp) / mh,  alt = alb + al,
//   p' = pb expm1(GAMMA (log1p(theta'/thb) [+ log1p(RVOVRD qv)]
//                        - log1p(al/alb))),  p = pb + p',
// which never differences the full geopotential nor subtracts pb from a
// full pressure.  The logarithms and the exponential run in double and p'
// is rounded once.
struct EosPert {
    real al, alt, pp, p;
};

__device__ __forceinline__ EosPert eos_pert(real thp, real thb, real qv,
                                            bool moist, real alb, real pb,
                                            real mup, real mh, real c1hk,
                                            real rdnwk, real dphp)
{
    EosPert o;
    o.al = -((alb * (c1hk * mup)) + rdnwk * dphp) / mh;
    o.alt = alb + o.al;
    double x = log1p((double)thp / (double)thb);
    if (moist) x = x + log1p((double)RVOVRD_F * (double)qv);
    const double y = log1p((double)o.al / (double)alb);
    o.pp = (real)((double)pb * expm1((double)GAMMA_F * (x - y)));
    o.p = pb + o.pp;
    return o;
}

// The EOS's specific volumes of layer k with column mass mu and
// geopotential thickness dph (reference diagnostics.cu:66-84):
// hypsometric_opt 1 (or 0) alt = -dph rdnw / C(mu), al = alt - alb;
// hypsometric_opt 2 al = dph / p_mid / log(p_bot / p_top_of_layer) - alb,
// alt = al + alb, with the layer's dry pressures from the hybrid
// coefficients c3/c4 and p_top.  The logarithm runs in double and is
// rounded once.  XO = false compiles the hypsometric_opt = 1 form alone
// (the step's plain variants, launched when no option needs the others).
template <bool XO = true>
__device__ __forceinline__ void eos_alpha(const Dom& d, int k, real mu,
                                          real dph, real albk, real c1k,
                                          real c2k, real rdnwk, real& alt,
                                          real& al)
{
    if (XO && d.hypsometric_opt == 2) {
        const real pt = (real)d.p_top;
        const real pfu = (d.c3f[k + 1] * mu + d.c4f[k + 1]) + pt;
        const real pfd = (d.c3f[k] * mu + d.c4f[k]) + pt;
        const real phm = (d.c3h[k] * mu + d.c4h[k]) + pt;
        const real lg = (real)log((double)(pfd / pfu));
        al = ((dph / phm) / lg) - albk;
        alt = al + albk;
    } else {
        alt = (-dph * rdnwk) / (c1k * mu + c2k);
        al = alt - albk;
    }
}

// Base-profile value: 1-D column (base3d = 0) or per-point 3-D field.
__device__ __forceinline__ real basev(const real* __restrict__ b, int k,
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
__device__ __forceinline__ real gld(const real* base, unsigned lev,
                                     unsigned colb)
{
    const char* b = reinterpret_cast<const char*>(base + lev);
    return *reinterpret_cast<const real*>(b + (uint64_t)colb);
}

__device__ __forceinline__ void gst(real* base, unsigned lev, unsigned colb,
                                    real x)
{
    char* b = reinterpret_cast<char*>(base + lev);
    *reinterpret_cast<real*>(b + (uint64_t)colb) = x;
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
typedef const __attribute__((address_space(4))) real kreal;

__device__ __forceinline__ kreal* kview(const real* p)
{
    return (kreal*)p;
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

// The same numbering over the first n workgroups of a launch whose later
// workgroups do other work (n a multiple of ARW_NXCD).
__device__ __forceinline__ int xcd_block_n(int n)
{
    const int b = (int)blockIdx.x;
    return (b % ARW_NXCD) * (n / ARW_NXCD) + b / ARW_NXCD;
}

static inline unsigned xcd_grid(unsigned nblocks)
{
    return (nblocks + ARW_NXCD - 1) / ARW_NXCD * ARW_NXCD;
}

// Pin a value as computed unconditionally.  A select between a cheap and
// an expensive operand (a division) is otherwise turned back into a
// branch by the code generator, splitting a level loop into blocks.
__device__ __forceinline__ void keep(real& x)
{
    asm volatile("" : "+v"(x));
}

// Compensated (Kahan) running sum of the precision form: s carries the
// sum, e the low-order part lost by the last addition.
struct Kahan {
    real s = R(0.0), e = R(0.0);
    __device__ __forceinline__ void add(real x)
    {
        const real y = x - e;
        const real t = s + y;
        e = (t - s) - y;
        s = t;
    }
};

// Column sums split over KG level groups: thread (tx, ty) of a (64, KG)
// block sums levels ty, ty + KG, ... of column blockIdx.x * 64 + tx (each
// level's loads coalesced over tx), then the groups meet in LDS.  Four
// times the waves of a thread-per-column loop, which at a few tens of
// thousands of columns leaves most SIMDs without a second wave to hide
// load latency behind.
constexpr int KG = 4;

template <int N>
__device__ __forceinline__ void group_sums(real (&v)[N])
{
    static_assert(KG == 4, "group_sums adds four level groups");
    __shared__ real sh[N][KG][64];
    for (int q = 0; q < N; ++q) sh[q][threadIdx.y][threadIdx.x] = v[q];
    __syncthreads();
    for (int q = 0; q < N; ++q)
        v[q] = (sh[q][0][threadIdx.x] + sh[q][1][threadIdx.x])
               + (sh[q][2][threadIdx.x] + sh[q][3][threadIdx.x]);
    __syncthreads();                 // sh may be reused by a later call
}
