// Dycore-step kernels for MI355X (gfx950) and the C-ABI that drives them.
//
// One RK3 split-explicit step (reference gpuwm/core/dycore.py:2161-2503)
// is issued as:
//   per stage   k_stage_col   (EOS, Omega, small-step init, implicit-w
//                              coefficients, reference mass divergence)
//               k_slow_tend   (advection + PGF + buoyancy + rhs_ph +
//                              constant-K diffusion, u''/v'' init)
//   per substep k_advance_uv  (u''/v'' forward step)
//               k_acoustic_col (mu''/Omega''/(mu theta)'' + implicit w''/phi''
//                               + linearised EOS, one thread per column)
//   per stage   k_finish_uv, k_finish_col (fold perturbations, EOS)
//   per step    k_set_w_surface
// Wavefront mapping: 64-wide waves run along x (the contiguous axis), so
// every level of every field is read as 256-byte coalesced rows.
#include "common.h"
#include "arwen_dycore.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <new>

typedef arw_domain Dom;

// Float32 scalars derived from the double-precision configuration exactly
// as the reference's host code derives them before a launch.
struct Sc {
    float rdx, rdy;          // float32(1/dx)
    float hrdx, hrdy;        // float32(0.5/dx)
    float qrdx, qrdy;        // float32(0.25/dx)
    float dxi2, dyi2;        // float32(1/dx**2)
    float kh, kv;
    float dtau, epssm, smdiv;
    float dampmag, zdamp;
    int top_lid, diff, damp;
};

static thread_local char g_err[512] = "";

static int32_t fail(int32_t code, const char* msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

static Sc make_sc(const arw_params* p, double dtau, int first)
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
    s.damp = p->damp_opt == 3;
    s.dampmag = s.damp ? (float)(dtau * p->dampcoef) : 0.0f;
    s.zdamp = (float)p->zdamp;
    s.top_lid = p->top_lid != 0;
    s.diff = (p->khdif > 0.0 || p->kvdif > 0.0);
    return s;
}

// ---------------------------------------------------------------------------
// Shared device pieces
// ---------------------------------------------------------------------------

__device__ __forceinline__ float mut_at(const Dom& d, int c)
{
    return d.mub2d[c] + d.mup[c];
}

// Coupled x mass flux at u face f (0..nx) of row j, level k:
// (c1h*<mu>_x + c2h)*u with <mu>_x = 0.5*(mu[f] + mu[f-1]) (periodic).
__device__ __forceinline__ float ru_at(const Dom& d, int k, int j, int f)
{
    const int nx = d.nx, ny = d.ny;
    int a = (f >= nx) ? f - nx : f;
    int b = (f == 0) ? nx - 1 : f - 1;
    float mux = 0.5f * (mut_at(d, j * nx + a) + mut_at(d, j * nx + b));
    return (d.c1h[k] * mux + d.c2h[k]) * d.u[iU(k, j, f, ny, nx)];
}

__device__ __forceinline__ float rv_at(const Dom& d, int k, int g, int i)
{
    const int nx = d.nx, ny = d.ny;
    int a = (g >= ny) ? g - ny : g;
    int b = (g == 0) ? ny - 1 : g - 1;
    float muy = 0.5f * (mut_at(d, a * nx + i) + mut_at(d, b * nx + i));
    return (d.c1h[k] * muy + d.c2h[k]) * d.v[iV(k, g, i, ny, nx)];
}

// Vertical face flux of a half-level field at w-level kf (reference
// advection.cu:189-201): zero at the boundaries, fnm/fnp 2nd order one
// face in, Omega-signed flux3 inside.
template <class Q>
__device__ __forceinline__ float zface_half(Q q, float vel, int kf, int nz,
                                            const float* __restrict__ fnm,
                                            const float* __restrict__ fnp)
{
    if (kf == 0 || kf == nz) return 0.0f;
    if (kf == 1 || kf == nz - 1)
        return vel * (fnm[kf] * q(kf) + fnp[kf] * q(kf - 1));
    return flux3(q(kf - 2), q(kf - 1), q(kf), q(kf + 1), vel);
}

__device__ __forceinline__ float div3(float dfx, float dfy, float dfz,
                                      float rdx, float rdy, float rdz)
{
    float t = -dfx * rdx;
    t = t - dfy * rdy;
    return t - dfz * rdz;
}

// ---------------------------------------------------------------------------
// Advection device functions (periodic, msf = 1)
// ---------------------------------------------------------------------------

// Scalar q at mass point (k,j,i); q(kk,jj,ii) accessor with wrapped
// indices; RU(k,j,f), RV(k,g,i), RW(kf,j,i) advecting fluxes.
template <class Q, class RU, class RV, class RW>
__device__ float adv_scalar(Q q, RU ru, RV rv, RW rw, int k, int j, int i,
                            int nz, int ny, int nx,
                            const float* __restrict__ rdnw,
                            const float* __restrict__ fnm,
                            const float* __restrict__ fnp,
                            float rdx, float rdy)
{
    float x[7], y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        x[o] = q(k, j, pmod(i + o - 3, nx));
        y[o] = q(k, pmod(j + o - 3, ny), i);
    }
    float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], ru(k, j, i));
    float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], ru(k, j, i + 1));
    float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], rv(k, j, i));
    float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], rv(k, j + 1, i));
    auto qz = [&](int kk) { return q(kk, j, i); };
    float fz0 = zface_half(qz, rw(k, j, i), k, nz, fnm, fnp);
    float fz1 = zface_half(qz, rw(k + 1, j, i), k + 1, nz, fnm, fnp);
    return div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, rdx, rdy, rdnw[k]);
}

// u momentum at u point (k, j, i), i = 0..nx (reference advection.cu:
// 572-611).  u(kk,jj,ii) reads the u array at a wrapped core column.
template <class U, class RU, class RV, class RW>
__device__ float adv_u(U u, RU ru, RV rv, RW rw, int k, int j, int i,
                       int nz, int ny, int nx,
                       const float* __restrict__ rdnw,
                       const float* __restrict__ fnm,
                       const float* __restrict__ fnp, float rdx, float rdy)
{
    int ic = (i >= nx) ? i - nx : i;
    int im = (ic == 0) ? nx - 1 : ic - 1;
    // x faces at mass centres im (left) and ic (right)
    float x[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) x[o] = u(k, j, pmod(ic + o - 3, nx));
    float vl = 0.5f * (ru(k, j, im) + ru(k, j, im + 1));
    float vr = 0.5f * (ru(k, j, ic) + ru(k, j, ic + 1));
    float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vl);
    float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], vr);
    float y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) y[o] = u(k, pmod(j + o - 3, ny), ic);
    float v0 = 0.5f * (rv(k, j, im) + rv(k, j, ic));
    float v1 = 0.5f * (rv(k, j + 1, im) + rv(k, j + 1, ic));
    float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
    float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
    auto qz = [&](int kk) { return u(kk, j, ic); };
    float w0 = 0.5f * (rw(k, j, im) + rw(k, j, ic));
    float w1 = 0.5f * (rw(k + 1, j, im) + rw(k + 1, j, ic));
    float fz0 = zface_half(qz, w0, k, nz, fnm, fnp);
    float fz1 = zface_half(qz, w1, k + 1, nz, fnm, fnp);
    return div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, rdx, rdy, rdnw[k]);
}

// v momentum at v point (k, j, i), j = 0..ny (reference advection.cu:
// 754-794).
template <class V, class RU, class RV, class RW>
__device__ float adv_v(V v, RU ru, RV rv, RW rw, int k, int j, int i,
                       int nz, int ny, int nx,
                       const float* __restrict__ rdnw,
                       const float* __restrict__ fnm,
                       const float* __restrict__ fnp, float rdx, float rdy)
{
    int jc = (j >= ny) ? j - ny : j;
    int jm = (jc == 0) ? ny - 1 : jc - 1;
    float x[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) x[o] = v(k, jc, pmod(i + o - 3, nx));
    float u0 = 0.5f * (ru(k, jm, i) + ru(k, jc, i));
    float u1 = 0.5f * (ru(k, jm, i + 1) + ru(k, jc, i + 1));
    float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0);
    float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1);
    float y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) y[o] = v(k, pmod(jc + o - 3, ny), i);
    float vl = 0.5f * (rv(k, jm, i) + rv(k, jm + 1, i));
    float vr = 0.5f * (rv(k, jc, i) + rv(k, jc + 1, i));
    float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], vl);
    float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], vr);
    auto qz = [&](int kk) { return v(kk, jc, i); };
    float w0 = 0.5f * (rw(k, jm, i) + rw(k, jc, i));
    float w1 = 0.5f * (rw(k + 1, jm, i) + rw(k + 1, jc, i));
    float fz0 = zface_half(qz, w0, k, nz, fnm, fnp);
    float fz1 = zface_half(qz, w1, k + 1, nz, fnm, fnp);
    return div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, rdx, rdy, rdnw[k]);
}

// w momentum at interior w level k = 1..nz-1 (reference advection.cu:
// 1016-1050).
template <class W, class RU, class RV, class RW>
__device__ float adv_w(W w, RU ru, RV rv, RW rw, int k, int j, int i,
                       int nz, int ny, int nx,
                       const float* __restrict__ rdn,
                       const float* __restrict__ fnm,
                       const float* __restrict__ fnp, float rdx, float rdy)
{
    float x[7], y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        x[o] = w(k, j, pmod(i + o - 3, nx));
        y[o] = w(k, pmod(j + o - 3, ny), i);
    }
    float u0 = fnm[k] * ru(k, j, i) + fnp[k] * ru(k - 1, j, i);
    float u1 = fnm[k] * ru(k, j, i + 1) + fnp[k] * ru(k - 1, j, i + 1);
    float v0 = fnm[k] * rv(k, j, i) + fnp[k] * rv(k - 1, j, i);
    float v1 = fnm[k] * rv(k, j + 1, i) + fnp[k] * rv(k - 1, j + 1, i);
    float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0);
    float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1);
    float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
    float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
    float fz[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        int m = k - 1 + s;
        float vel = 0.5f * (rw(m, j, i) + rw(m + 1, j, i));
        if (m == 0 || m == nz - 1)
            fz[s] = 0.5f * vel * (w(m + 1, j, i) + w(m, j, i));
        else
            fz[s] = flux3(w(m - 1, j, i), w(m, j, i), w(m + 1, j, i),
                          w(m + 2, j, i), vel);
    }
    return div3(fx1 - fx0, fy1 - fy0, fz[1] - fz[0], rdx, rdy, rdn[k]);
}

// ---------------------------------------------------------------------------
// Standalone advection kernels (reference kernel signatures)
// ---------------------------------------------------------------------------

struct Arr3 {
    const float* p;
    int ny, nx;      // row count and row width of the array
    __device__ float operator()(int k, int j, int i) const
    {
        return p[(k * ny + j) * nx + i];
    }
};

__global__ void k_flux_div(int which, const float* __restrict__ q,
                           const float* __restrict__ ru,
                           const float* __restrict__ rv,
                           const float* __restrict__ rw,
                           float* __restrict__ tend,
                           const float* __restrict__ rdz,
                           const float* __restrict__ fnm,
                           const float* __restrict__ fnp, float rdx,
                           float rdy, int nz, int ny, int nx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    Arr3 RU{ru, ny, nx + 1}, RV{rv, ny + 1, nx}, RW{rw, ny, nx};
    if (which == 0) {
        if (i >= nx || j >= ny || k >= nz) return;
        Arr3 Q{q, ny, nx};
        tend[iM(k, j, i, ny, nx)] += adv_scalar(Q, RU, RV, RW, k, j, i, nz,
                                                ny, nx, rdz, fnm, fnp, rdx,
                                                rdy);
    } else if (which == 1) {
        if (i > nx || j >= ny || k >= nz) return;
        Arr3 Q{q, ny, nx + 1};
        tend[iU(k, j, i, ny, nx)] += adv_u(Q, RU, RV, RW, k, j, i, nz, ny,
                                           nx, rdz, fnm, fnp, rdx, rdy);
    } else if (which == 2) {
        if (i >= nx || j > ny || k >= nz) return;
        Arr3 Q{q, ny + 1, nx};
        tend[iV(k, j, i, ny, nx)] += adv_v(Q, RU, RV, RW, k, j, i, nz, ny,
                                           nx, rdz, fnm, fnp, rdx, rdy);
    } else {
        if (i >= nx || j >= ny || k > nz) return;
        if (k == 0 || k == nz) return;
        Arr3 Q{q, ny, nx};
        tend[iM(k, j, i, ny, nx)] += adv_w(Q, RU, RV, RW, k, j, i, nz, ny,
                                           nx, rdz, fnm, fnp, rdx, rdy);
    }
}

// ---------------------------------------------------------------------------
// Pressure-gradient pieces
// ---------------------------------------------------------------------------

__device__ __forceinline__ float pp_at(const Dom& d, int k, int c, int st)
{
    return d.p[k * st + c] - basev(d.pb, k, c, st, d.base3d);
}

// Perturbation pressure at full level kf of column c (WRF dpn).
__device__ __forceinline__ float dpn_at(const Dom& d, int kf, int c, int st,
                                        int top_lid)
{
    const int nz = d.nz;
    if (kf == 0)
        return (d.cf1 * pp_at(d, 0, c, st) + d.cf2 * pp_at(d, 1, c, st))
               + d.cf3 * pp_at(d, 2, c, st);
    if (kf == nz) {
        if (!top_lid) return 0.0f;
        return d.cfn * pp_at(d, nz - 1, c, st)
               + d.cfn1 * pp_at(d, nz - 2, c, st);
    }
    return d.fnm[kf] * pp_at(d, kf, c, st)
           + d.fnp[kf] * pp_at(d, kf - 1, c, st);
}

__device__ __forceinline__ float phh_at(const Dom& d, int k, int c, int st)
{
    if (d.base3d)
        return 0.5f * ((d.phb[k * st + c] + d.phb[(k + 1) * st + c])
                       + (d.php[k * st + c] + d.php[(k + 1) * st + c]));
    return 0.5f * (d.php[k * st + c] + d.php[(k + 1) * st + c]);
}

// Large-step horizontal PGF on the face between columns B (left/south)
// and A (reference dycore.cu:98-157 pgf_face).
__device__ float pgf_face(const Dom& d, int k, int cA, int cB, int st,
                          float rd, float hrd, int top_lid)
{
    float muf = 0.5f * (mut_at(d, cA) + mut_at(d, cB));
    float dmu = 0.5f * (d.mup[cA] + d.mup[cB]);
    float layer = d.c1h[k] * muf + d.c2h[k];
    int a = k * st + cA, b = k * st + cB;
    float bracket = (d.php[a + st] - d.php[b + st]) + (d.php[a] - d.php[b]);
    bracket = bracket + (d.alt[a] + d.alt[b])
                            * (pp_at(d, k, cA, st) - pp_at(d, k, cB, st));
    if (d.base3d)
        bracket = bracket + (d.al[a] + d.al[b]) * (d.pb[a] - d.pb[b]);
    else
        bracket = bracket + (d.al[a] + d.al[b]) * 0.0f;
    float left = (hrd * layer) * bracket;
    float dphp = phh_at(d, k, cA, st) - phh_at(d, k, cB, st);
    float hi = 0.5f * (dpn_at(d, k + 1, cA, st, top_lid)
                       + dpn_at(d, k + 1, cB, st, top_lid));
    float lo = 0.5f * (dpn_at(d, k, cA, st, top_lid)
                       + dpn_at(d, k, cB, st, top_lid));
    float vertical = d.rdnw[k] * (hi - lo) - d.c1h[k] * dmu;
    return left + (rd * dphp) * vertical;
}

// Constant-K Laplacian at (k, j, i) of a field with accessor f over its
// periodic core (ny, nx) (reference diffusion.cu:22-51).
template <class F>
__device__ float diff2(F f, int k, int j, int i, int nlev, int ny, int nx,
                       int wstag, const Sc& s,
                       const float* __restrict__ rdzf,
                       const float* __restrict__ rdzc)
{
    if (wstag && (k == 0 || k == nlev - 1)) return 0.0f;
    float fc = f(k, j, i);
    float lap = s.kh * (s.dxi2 * ((f(k, j, pmod(i + 1, nx)) - 2.0f * fc)
                                  + f(k, j, pmod(i - 1, nx)))
                        + s.dyi2 * ((f(k, pmod(j + 1, ny), i) - 2.0f * fc)
                                    + f(k, pmod(j - 1, ny), i)));
    float fup = (k < nlev - 1) ? (f(k + 1, j, i) - fc) * rdzf[k] : 0.0f;
    float fdn = (k > 0) ? (fc - f(k - 1, j, i)) * rdzf[k - 1] : 0.0f;
    return lap + (s.kv * (fup - fdn)) * rdzc[k];
}

// ---------------------------------------------------------------------------
// Stage column kernel: EOS, Omega, small-step init, coefficients
// ---------------------------------------------------------------------------

__device__ __forceinline__ void eos_level(const Dom& d, int k, int c, int st,
                                          float mu)
{
    float th = basev(d.thb, k, c, st, d.base3d) + d.thp[k * st + c];
    float ph_hi = basev(d.phb, k + 1, c, st, d.base3d)
                  + d.php[(k + 1) * st + c];
    float ph_lo = basev(d.phb, k, c, st, d.base3d) + d.php[k * st + c];
    float dph = ph_hi - ph_lo;
    float alt = (-dph * d.rdnw[k]) / (d.c1h[k] * mu + d.c2h[k]);
    d.alt[k * st + c] = alt;
    d.al[k * st + c] = alt - basev(d.alb, k, c, st, d.base3d);
    d.p[k * st + c] = P0_F * pow_gamma((RD_F * th) / (P0_F * alt));
}

__global__ void k_eos(Dom d)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int st = d.ny * d.nx;
    if (c >= st) return;
    float mu = mut_at(d, c);
    for (int k = 0; k < d.nz; ++k) eos_level(d, k, c, st, mu);
}

// Layer divergence of the coupled reference fluxes at level k of column
// (j, i) (reference dycore.py:117-119).
__device__ __forceinline__ float divv_at(const Dom& d, const Sc& s, int k,
                                         int j, int i, float muxw,
                                         float muxe, float muys, float muyn)
{
    const int nx = d.nx, ny = d.ny;
    float ruw = (d.c1h[k] * muxw + d.c2h[k]) * d.u[iU(k, j, i, ny, nx)];
    float rue = (d.c1h[k] * muxe + d.c2h[k]) * d.u[iU(k, j, i + 1, ny, nx)];
    float rvs = (d.c1h[k] * muys + d.c2h[k]) * d.v[iV(k, j, i, ny, nx)];
    float rvn = (d.c1h[k] * muyn + d.c2h[k]) * d.v[iV(k, j + 1, i, ny, nx)];
    return d.dnw[k] * (s.rdx * (rue - ruw) + s.rdy * (rvn - rvs));
}

__global__ void k_stage_col(Dom d, Sc s, int do_eos, int do_init,
                            int do_coef)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    int j = c / nx, i = c - j * nx;
    int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    float ms = mut_at(d, c);
    if (do_eos)
        for (int k = 0; k < nz; ++k) eos_level(d, k, c, st, ms);

    // Omega (WRF calc_ww_cp, reference dycore.py:102-131)
    float mw = mut_at(d, j * nx + im), me = mut_at(d, j * nx + ip);
    float mS = mut_at(d, jm * nx + i), mN = mut_at(d, jp * nx + i);
    float muxw = 0.5f * (ms + mw), muxe = 0.5f * (me + ms);
    float muys = 0.5f * (ms + mS), muyn = 0.5f * (mN + ms);
    float dmdt = divv_at(d, s, 0, j, i, muxw, muxe, muys, muyn);
    for (int k = 1; k < nz; ++k)
        dmdt = dmdt + divv_at(d, s, k, j, i, muxw, muxe, muys, muyn);
    d.ww[c] = 0.0f;
    float run = 0.0f;
    for (int k = 0; k < nz - 1; ++k) {
        float inc = (d.c1h[k] * d.dnw[k]) * dmdt
                    + divv_at(d, s, k, j, i, muxw, muxe, muys, muyn);
        run = run + inc;
        d.ww[(k + 1) * st + c] = -run;
    }
    d.ww[nz * st + c] = 0.0f;
    d.rmu_t[c] = 0.0f;

    if (do_init) {
        // WRF small_step_prep + calc_p_rho (reference dycore.cu:1007-1085)
        float mt = d.mub2d[c] + d.mup0[c];
        float mupp = d.mup0[c] - d.mup[c];
        d.mu_pp[c] = mupp;
        for (int k = 0; k <= nz; ++k) {
            int f = k * st + c;
            d.w_pp[f] = (d.c1f[k] * mt + d.c2f[k]) * d.w0[f]
                        - (d.c1f[k] * ms + d.c2f[k]) * d.w[f];
            d.ph_pp[f] = d.php0[f] - d.php[f];
        }
        for (int k = 0; k < nz; ++k) {
            int h = k * st + c;
            float cht = d.c1h[k] * mt + d.c2h[k];
            float chs = d.c1h[k] * ms + d.c2h[k];
            float thb = basev(d.thb, k, c, st, d.base3d);
            float tht = thb + d.thp0[h];
            float ths = thb + d.thp[h];
            float thpp = cht * tht - chs * ths;
            d.th_pp[h] = thpp;
            float alt = d.alt[h];
            float c2a = (GAMMA_F * d.p[h]) / alt;
            float al = -((alt * (d.c1h[k] * mupp))
                         + d.rdnw[k] * (d.ph_pp[h + st] - d.ph_pp[h])) / cht;
            d.al_pp[h] = al;
            float p5 = (alt * (thpp - (d.c1h[k] * mupp) * ths)) / (cht * ths);
            float pn = c2a * (p5 - al);
            d.p_pp[h] = pn;
            d.p_pp_old[h] = pn;
        }
    }

    if (do_coef) {
        // calc_coef_w (reference acoustic.cu:675-759) and the substeps'
        // constant reference mass divergence (acoustic.cu:396-405).
        float cof = 0.5f * s.dtau * G_F * (1.0f + s.epssm);
        cof = cof * cof;
        for (int k = 0; k < nz; ++k)
            d.c2a[k * st + c] = GAMMA_F * d.p[k * st + c] / d.alt[k * st + c];
        d.a[c] = 0.0f;
        d.alpha[c] = 0.0f;
        d.gam[c] = 0.0f;
        d.a[st + c] = 0.0f;
        for (int k = 2; k < nz; ++k) {
            float chm = d.c1h[k - 1] * ms + d.c2h[k - 1];
            float cfm = d.c1f[k - 1] * ms + d.c2f[k - 1];
            d.a[k * st + c] = -cof * d.rdn[k] * d.rdnw[k - 1]
                              * d.c2a[(k - 1) * st + c] / (chm * cfm);
        }
        if (s.top_lid) {
            d.a[nz * st + c] = 0.0f;
        } else {
            float chm = d.c1h[nz - 1] * ms + d.c2h[nz - 1];
            float cfm = d.c1f[nz - 1] * ms + d.c2f[nz - 1];
            d.a[nz * st + c] = -2.0f * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                               * d.c2a[(nz - 1) * st + c] / (chm * cfm);
        }
        float gprev = 0.0f;
        for (int k = 1; k < nz; ++k) {
            float chm_k = d.c1h[k] * ms + d.c2h[k];
            float chm_km = d.c1h[k - 1] * ms + d.c2h[k - 1];
            float cfm_k = d.c1f[k] * ms + d.c2f[k];
            float cfm_kp = d.c1f[k + 1] * ms + d.c2f[k + 1];
            float c2k = d.c2a[k * st + c], c2km = d.c2a[(k - 1) * st + c];
            float b = 1.0f + cof * d.rdn[k]
                                 * (d.rdnw[k] * c2k / (chm_k * cfm_k)
                                    + d.rdnw[k - 1] * c2km / (chm_km * cfm_k));
            float cc = -cof * d.rdn[k] * d.rdnw[k] * c2k / (chm_k * cfm_kp);
            float al = 1.0f / (b - d.a[k * st + c] * gprev);
            d.alpha[k * st + c] = al;
            gprev = cc * al;
            d.gam[k * st + c] = gprev;
        }
        float chm = d.c1h[nz - 1] * ms + d.c2h[nz - 1];
        float cfn = d.c1f[nz] * ms + d.c2f[nz];
        float b_top = 1.0f + 2.0f * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                                 * d.c2a[(nz - 1) * st + c] / (chm * cfn);
        d.alpha[nz * st + c] = 1.0f / (b_top - d.a[nz * st + c] * gprev);
        d.gam[nz * st + c] = 0.0f;

        // sum_k dnw*div(C(mu_face) u_t*) over the column
        float rf = 0.0f;
        for (int k = 0; k < nz; ++k) {
            float dvr = s.rdx * ((d.c1h[k] * muxe + d.c2h[k])
                                     * d.u[iU(k, j, i + 1, ny, nx)]
                                 - (d.c1h[k] * muxw + d.c2h[k])
                                       * d.u[iU(k, j, i, ny, nx)])
                        + s.rdy * ((d.c1h[k] * muyn + d.c2h[k])
                                       * d.v[iV(k, j + 1, i, ny, nx)]
                                   - (d.c1h[k] * muys + d.c2h[k])
                                         * d.v[iV(k, j, i, ny, nx)]);
            rf = rf + d.dnw[k] * dvr;
        }
        d.dmdt_rf[c] = rf;
    }
}

// ---------------------------------------------------------------------------
// Slow tendencies (one thread per point of the union staggered grid)
// ---------------------------------------------------------------------------

__global__ void k_slow_tend(Dom d, Sc s, int do_init_uv)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;

    auto RU = [&](int kk, int jj, int ff) { return ru_at(d, kk, jj, ff); };
    auto RV = [&](int kk, int gg, int ii) { return rv_at(d, kk, gg, ii); };
    auto RW = [&](int kk, int jj, int ii) { return d.ww[iM(kk, jj, ii, ny, nx)]; };

    if (k < nz && j < ny) {                       // u point
        auto U = [&](int kk, int jj, int ii) { return d.u[iU(kk, jj, ii, ny, nx)]; };
        float t = adv_u(U, RU, RV, RW, k, j, i, nz, ny, nx, d.rdnw, d.fnm,
                        d.fnp, s.rdx, s.rdy);
        int ic = (i >= nx) ? i - nx : i;
        int ib = (ic == 0) ? nx - 1 : ic - 1;
        t = t - pgf_face(d, k, j * nx + ic, j * nx + ib, st, s.rdx, s.hrdx,
                         s.top_lid);
        float mA = mut_at(d, j * nx + ic), mB = mut_at(d, j * nx + ib);
        if (s.diff) {
            float lap = diff2(U, k, j, ic, nz, ny, nx, 0, s, d.rdzf_h, d.rdzc_h);
            float mux = 0.5f * (mA + mB);
            t = t + (d.c1h[k] * mux + d.c2h[k]) * lap;
        }
        int ix = iU(k, j, i, ny, nx);
        d.ru_t[ix] = t;
        if (do_init_uv) {
            float mtf = 0.5f * ((d.mub2d[j * nx + ic] + d.mup0[j * nx + ic])
                                + (d.mub2d[j * nx + ib] + d.mup0[j * nx + ib]));
            float msf = 0.5f * (mA + mB);
            d.u_pp[ix] = (d.c1h[k] * mtf + d.c2h[k]) * d.u0[ix]
                         - (d.c1h[k] * msf + d.c2h[k]) * d.u[ix];
        }
    }
    if (k < nz && i < nx) {                       // v point
        auto V = [&](int kk, int jj, int ii) { return d.v[iV(kk, jj, ii, ny, nx)]; };
        float t = adv_v(V, RU, RV, RW, k, j, i, nz, ny, nx, d.rdnw, d.fnm,
                        d.fnp, s.rdx, s.rdy);
        int jc = (j >= ny) ? j - ny : j;
        int jb = (jc == 0) ? ny - 1 : jc - 1;
        t = t - pgf_face(d, k, jc * nx + i, jb * nx + i, st, s.rdy, s.hrdy,
                         s.top_lid);
        float mA = mut_at(d, jc * nx + i), mB = mut_at(d, jb * nx + i);
        if (s.diff) {
            float lap = diff2(V, k, jc, i, nz, ny, nx, 0, s, d.rdzf_h, d.rdzc_h);
            float muy = 0.5f * (mA + mB);
            t = t + (d.c1h[k] * muy + d.c2h[k]) * lap;
        }
        int ix = iV(k, j, i, ny, nx);
        d.rv_t[ix] = t;
        if (do_init_uv) {
            float mtf = 0.5f * ((d.mub2d[jc * nx + i] + d.mup0[jc * nx + i])
                                + (d.mub2d[jb * nx + i] + d.mup0[jb * nx + i]));
            float msf = 0.5f * (mA + mB);
            d.v_pp[ix] = (d.c1h[k] * mtf + d.c2h[k]) * d.v0[ix]
                         - (d.c1h[k] * msf + d.c2h[k]) * d.v[ix];
        }
    }
    if (i < nx && j < ny) {                       // w point / mass point
        int c = j * nx + i;
        float mu = mut_at(d, c);
        auto W = [&](int kk, int jj, int ii) { return d.w[iM(kk, jj, ii, ny, nx)]; };
        // rw_t: advection (interior levels) + buoyancy + diffusion
        float t = 0.0f;
        if (k >= 1 && k < nz)
            t = adv_w(W, RU, RV, RW, k, j, i, nz, ny, nx, d.rdn, d.fnm,
                      d.fnp, s.rdx, s.rdy);
        if (k >= 1 && k < nz) {
            float term = d.rdn[k] * (pp_at(d, k, c, st) - pp_at(d, k - 1, c, st))
                         - d.c1f[k] * d.mup[c];
            t = t + G_F * term;
        } else if (k == nz) {
            float pressure = ((1.0f * 2.0f) * d.rdnw[nz - 1])
                             * (0.0f - pp_at(d, nz - 1, c, st));
            float term = pressure - d.c1f[nz] * d.mup[c];
            t = t + G_F * term;
        }
        if (s.diff) {
            float lap = diff2(W, k, j, i, nz + 1, ny, nx, 1, s, d.rdzf_w, d.rdzc_w);
            t = t + (d.c1f[k] * mu + d.c2f[k]) * lap;
        }
        d.rw_t[k * st + c] = t;

        // rph_t: WRF rhs_ph (reference dycore.cu:541-667, order 2)
        float r = 0.0f;
        if (k >= 1) {
            int top = (k == nz);
            auto ph = [&](int kk, int cc) {
                return basev(d.phb, kk, cc, st, d.base3d) + d.php[kk * st + cc];
            };
            if (!top) {
                float wd_lo = ((0.5f * (d.ww[k * st + c] + d.ww[(k - 1) * st + c]))
                               * d.rdnw[k - 1]) * (ph(k, c) - ph(k - 1, c));
                float wd_hi = ((0.5f * (d.ww[(k + 1) * st + c] + d.ww[k * st + c]))
                               * d.rdnw[k]) * (ph(k + 1, c) - ph(k, c));
                float omega = d.fnm[k] * wd_hi + d.fnp[k] * wd_lo;
                r = r - omega;
            }
            float mass = d.c1f[k] * mu + d.c2f[k];
            r = r + mass * (G_F * d.w[k * st + c]);
            float qx = top ? 2.0f * s.qrdx : s.qrdx;
            float qy = top ? 2.0f * s.qrdy : s.qrdy;
            int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
            int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
            auto fcx = [&](int f) {            // f in 0..nx-1 (wrapped face)
                int b = (f == 0) ? nx - 1 : f - 1;
                float avg = 0.5f * (mut_at(d, j * nx + f) + mut_at(d, j * nx + b));
                float m = d.c1f[k] * avg + d.c2f[k];
                float wind = top ? d.cfn * d.u[iU(k - 1, j, f, ny, nx)]
                                       + d.cfn1 * d.u[iU(k - 2, j, f, ny, nx)]
                                 : d.u[iU(k, j, f, ny, nx)] + d.u[iU(k - 1, j, f, ny, nx)];
                return m * wind;
            };
            auto fcy = [&](int g) {
                int b = (g == 0) ? ny - 1 : g - 1;
                float avg = 0.5f * (mut_at(d, g * nx + i) + mut_at(d, b * nx + i));
                float m = d.c1f[k] * avg + d.c2f[k];
                float wind = top ? d.cfn * d.v[iV(k - 1, g, i, ny, nx)]
                                       + d.cfn1 * d.v[iV(k - 2, g, i, ny, nx)]
                                 : d.v[iV(k, g, i, ny, nx)] + d.v[iV(k - 1, g, i, ny, nx)];
                return m * wind;
            };
            float p0 = ph(k, c);
            float fx = fcx(i) * (p0 - ph(k, j * nx + im));
            float fxp = fcx(ip) * (ph(k, j * nx + ip) - p0);
            r = r - qx * (fxp + fx);
            float fy = fcy(j) * (p0 - ph(k, jm * nx + i));
            float fyp = fcy(jp) * (ph(k, jp * nx + i) - p0);
            r = r - qy * (fyp + fy);
        }
        d.rph_t[k * st + c] = r;

        if (k < nz) {                             // rth_t
            auto Q = [&](int kk, int jj, int ii) {
                int cc = jj * nx + ii;
                return basev(d.thb, kk, cc, st, d.base3d) + d.thp[kk * st + cc];
            };
            float t2 = adv_scalar(Q, RU, RV, RW, k, j, i, nz, ny, nx, d.rdnw,
                                  d.fnm, d.fnp, s.rdx, s.rdy);
            if (s.diff) {
                auto T = [&](int kk, int jj, int ii) { return d.thp[iM(kk, jj, ii, ny, nx)]; };
                float lap = diff2(T, k, j, i, nz, ny, nx, 0, s, d.rdzf_h, d.rdzc_h);
                t2 = t2 + (d.c1h[k] * mu + d.c2h[k]) * lap;
            }
            d.rth_t[k * st + c] = t2;
        }
    }
}

// ---------------------------------------------------------------------------
// Acoustic substep
// ---------------------------------------------------------------------------

__device__ __forceinline__ float pdmp(const float* __restrict__ p,
                                      const float* __restrict__ pold, int ix,
                                      float smdiv)
{
    return p[ix] + smdiv * (p[ix] - pold[ix]);
}

// Face-averaged damped p'' at full level kf (reference acoustic.cu:163-190).
__device__ __forceinline__ float dpn_face(const Dom& d, const Sc& s, int kf,
                                          int cA, int cB, int st)
{
    const int nz = d.nz;
    auto S = [&](int kk) {
        return pdmp(d.p_pp, d.p_pp_old, kk * st + cA, s.smdiv)
               + pdmp(d.p_pp, d.p_pp_old, kk * st + cB, s.smdiv);
    };
    if (kf == 0)
        return 0.5f * ((d.cf1 * S(0) + d.cf2 * S(1)) + d.cf3 * S(2));
    if (kf == nz) {
        if (!s.top_lid) return 0.0f;
        return 0.5f * ((d.cf1 * S(nz - 1) + d.cf2 * S(nz - 2))
                       + d.cf3 * S(nz - 3));
    }
    return 0.5f * (d.fnm[kf] * S(kf) + d.fnp[kf] * S(kf - 1));
}

// Acoustic horizontal pressure gradient dpxy (reference acoustic.cu:
// 205-237 pgrad_face).
__device__ float pgrad_face(const Dom& d, const Sc& s, int k, int cA, int cB,
                            int st, float rd)
{
    float muf = 0.5f * (mut_at(d, cB) + mut_at(d, cA));
    int a = k * st + cA, b = k * st + cB;
    float dph = (d.ph_pp[a + st] - d.ph_pp[b + st]) + (d.ph_pp[a] - d.ph_pp[b]);
    float dpe = pdmp(d.p_pp, d.p_pp_old, a, s.smdiv)
                - pdmp(d.p_pp, d.p_pp_old, b, s.smdiv);
    float brk = dph + (d.alt[a] + d.alt[b]) * dpe;
    if (d.base3d) brk = brk + (d.al_pp[a] + d.al_pp[b]) * (d.pb[a] - d.pb[b]);
    float dpxy = ((0.5f * rd) * (d.c1h[k] * muf + d.c2h[k])) * brk;
    float up = dpn_face(d, s, k + 1, cA, cB, st);
    float dn = dpn_face(d, s, k, cA, cB, st);
    float phsA = d.php[a] + d.php[a + st], phsB = d.php[b] + d.php[b + st];
    float dphp = 0.5f * (phsA - phsB);
    if (d.base3d) {
        float pbA = d.phb[a] + d.phb[a + st], pbB = d.phb[b] + d.phb[b + st];
        dphp = dphp + 0.5f * (pbA - pbB);
    }
    return dpxy + (rd * dphp) * (d.rdnw[k] * (up - dn)
                                 - (0.5f * d.c1h[k]) * (d.mu_pp[cB] + d.mu_pp[cA]));
}

__global__ void k_advance_uv(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > nx || j > ny || k >= nz) return;
    if (j < ny) {
        int ic = (i >= nx) ? i - nx : i;
        int ib = (ic == 0) ? nx - 1 : ic - 1;
        float dpxy = pgrad_face(d, s, k, j * nx + ic, j * nx + ib, st, s.rdx);
        int ix = iU(k, j, i, ny, nx);
        d.u_pp[ix] = d.u_pp[ix] + s.dtau * (d.ru_t[ix] - dpxy);
    }
    if (i < nx) {
        int jc = (j >= ny) ? j - ny : j;
        int jb = (jc == 0) ? ny - 1 : jc - 1;
        float dpxy = pgrad_face(d, s, k, jc * nx + i, jb * nx + i, st, s.rdy);
        int ix = iV(k, j, i, ny, nx);
        d.v_pp[ix] = d.v_pp[ix] + s.dtau * (d.rv_t[ix] - dpxy);
    }
}

// mu''/Omega''/(mu theta)'' forward step (reference acoustic.cu:329-456),
// then the implicit w''/phi'' solve with the kinematic lower boundary,
// damp_opt=3 damper and the fused linearised EOS (acoustic.cu:774-982).
// One thread per column; the recurrences run in registers with sliding
// windows; w'' holds the forward-eliminated values and ph'' the phi''
// right-hand side between the upward and downward sweeps.  pout receives
// the new p''.
__global__ void k_acoustic_col(Dom d, Sc s, float* __restrict__ pout)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    int j = c / nx, i = c - j * nx;
    int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    const int b3 = d.base3d;
    const float eps = s.epssm, dtau = s.dtau;
    const float one_p = 1.0f + eps, one_m = 1.0f - eps;

    auto dvp_at = [&](int k) {
        return s.rdx * (d.u_pp[iU(k, j, i + 1, ny, nx)] - d.u_pp[iU(k, j, i, ny, nx)])
               + s.rdy * (d.v_pp[iV(k, j + 1, i, ny, nx)] - d.v_pp[iV(k, j, i, ny, nx)]);
    };
    float mut = mut_at(d, c);
    float mupp_old = d.mu_pp[c];
    float dmdt_pp = 0.0f;
    for (int k = 0; k < nz; ++k) dmdt_pp = dmdt_pp + d.dnw[k] * dvp_at(k);
    float rmu = d.rmu_t[c];
    float mupp = mupp_old + dtau * ((dmdt_pp + d.dmdt_rf[c]) + rmu);
    d.mu_pp[c] = mupp;
    float muts = mut + mupp;
    float muave = 0.5f * (one_p * mupp + one_m * mupp_old);

    auto thr = [&](int kk, int cc) {
        return basev(d.thb, kk, cc, st, b3) + d.thp[kk * st + cc];
    };
    auto phs = [&](int kk) { return basev(d.phb, kk, c, st, b3) + d.php[kk * st + c]; };

    // kinematic lower boundary on w'' from the updated u''/v''
    float w0;
    {
        auto lw = [&](const float* a, int ix0, int ix1, int ix2) {
            return (d.cf1 * a[ix0] + d.cf2 * a[ix1]) + d.cf3 * a[ix2];
        };
        float ue = lw(d.u_pp, iU(0, j, i + 1, ny, nx), iU(1, j, i + 1, ny, nx),
                      iU(2, j, i + 1, ny, nx));
        float uw = lw(d.u_pp, iU(0, j, i, ny, nx), iU(1, j, i, ny, nx),
                      iU(2, j, i, ny, nx));
        float vn = lw(d.v_pp, iV(0, j + 1, i, ny, nx), iV(1, j + 1, i, ny, nx),
                      iV(2, j + 1, i, ny, nx));
        float vs = lw(d.v_pp, iV(0, j, i, ny, nx), iV(1, j, i, ny, nx),
                      iV(2, j, i, ny, nx));
        float h = d.ht[c];
        w0 = (0.5f * s.rdy) * ((d.ht[jp * nx + i] - h) * vn + (h - d.ht[jm * nx + i]) * vs)
             + (0.5f * s.rdx) * ((d.ht[j * nx + ip] - h) * ue + (h - d.ht[j * nx + im]) * uw);
        d.w_pp[c] = w0;
    }

    // upward sweep
    d.ww_pp[c] = 0.0f;
    float ww_k = 0.0f;            // Omega'' at level m
    float wwk = 0.0f;             // running recurrence value
    float wdtn_k = 0.0f;          // vertical theta flux at level m
    float th_m = thr(0, c);       // theta_t* at level m
    float ph_m = phs(0);          // phi_t* at level m
    float wd_prev = 0.0f;         // wd at m-1
    // windows (index offsets relative to the row being formed)
    float rhs_a = 0.0f, rhs_b = 0.0f;          // rhs[m-2], rhs[m-1]
    float pho_a = 0.0f, pho_b = d.ph_pp[c];    // ph''old[m-2], [m-1]
    float t2_a = 0.0f, t2_b = 0.0f;            // t2[m-2], t2[m-1]
    float c2a_a = 0.0f, c2a_b = 0.0f;          // c2a at m-2, m-1
    float alt_a = 0.0f, alt_b = 0.0f;
    float chm_a = 0.0f, chm_b = 0.0f;
    float wold_b = 0.0f;                       // w''old[m-1]
    float wf_prev = w0;                        // forward-eliminated w'[r-1]
    const float ghalf_eps = (0.5f * G_F) * one_m;
    const float c_force = (0.5f * dtau) * G_F;
    const float c_buoy = dtau * G_F;

    auto force_row = [&](int r, float rhs_dn, float rhs_c, float rhs_up,
                         float ph_dn, float ph_c, float ph_up, float wold,
                         float t2dn, float t2c, float c2dn, float c2c,
                         float altdn, float altc, float chmdn, float chmc) {
        float dph_up = one_p * (rhs_up - rhs_c) + one_m * (ph_up - ph_c);
        float dph_dn = one_p * (rhs_c - rhs_dn) + one_m * (ph_c - ph_dn);
        float t1 = dtau * d.rw_t[r * st + c]
                   + (c_force * d.rdn[r])
                         * ((c2c * d.rdnw[r]) / chmc * dph_up
                            - (c2dn * d.rdnw[r - 1]) / chmdn * dph_dn);
        float t3 = c_buoy * (d.rdn[r] * ((c2c * altc) * t2c - (c2dn * altdn) * t2dn)
                             - d.c1f[r] * muave);
        return wold + (t1 + t3);
    };

    for (int m = 0; m < nz; ++m) {
        int h = m * st + c;
        float dvp = dvp_at(m);
        float ww_kp1;
        if (m + 1 < nz) {
            wwk = wwk - d.dnw[m] * (d.c1h[m] * (dmdt_pp + rmu) + dvp);
            ww_kp1 = wwk;
        } else {
            ww_kp1 = 0.0f;
        }
        d.ww_pp[(m + 1) * st + c] = ww_kp1;
        float th_up = (m + 1 < nz) ? thr(m + 1, c) : 0.0f;
        float wdtn_kp1 = (m + 1 < nz)
            ? ww_kp1 * (d.fnm[m + 1] * th_up + d.fnp[m + 1] * th_m) : 0.0f;
        // (mu theta)'' update
        float t_c = th_m;
        float hx = (0.5f * s.rdx) * (d.u_pp[iU(m, j, i + 1, ny, nx)] * (thr(m, j * nx + ip) + t_c)
                                     - d.u_pp[iU(m, j, i, ny, nx)] * (t_c + thr(m, j * nx + im)));
        float hy = (0.5f * s.rdy) * (d.v_pp[iV(m, j + 1, i, ny, nx)] * (thr(m, jp * nx + i) + t_c)
                                     - d.v_pp[iV(m, j, i, ny, nx)] * (t_c + thr(m, jm * nx + i)));
        float tho = d.th_pp[h];
        float thn = tho + dtau * (d.rth_t[h] - ((hx + hy) + d.rdnw[m] * (wdtn_kp1 - wdtn_k)));
        d.th_pp[h] = thn;
        float chs = d.c1h[m] * muts + d.c2h[m];
        float t2 = (0.5f * (one_p * thn + one_m * tho)) / (chs * t_c);
        // Omega''*d(phi_t*) at mass level m
        float ph_up = phs(m + 1);
        float wd = ((0.5f * (ww_kp1 + ww_k)) * d.rdnw[m]) * (ph_up - ph_m);
        float c2a = d.c2a[h], alt = d.alt[h];
        float chm = d.c1h[m] * mut + d.c2h[m];
        // rhs at full level m (m >= 1)
        float rhs_c = 0.0f, pho_c = pho_b, wold_c = 0.0f;
        if (m >= 1) {
            int f = m * st + c;
            pho_c = d.ph_pp[f];
            wold_c = d.w_pp[f];
            float r = dtau * (d.rph_t[f] + ghalf_eps * wold_c);
            r = r - dtau * (d.fnm[m] * wd + d.fnp[m] * wd_prev);
            float cfm = d.c1f[m] * mut + d.c2f[m];
            rhs_c = pho_c + r / cfm;
            d.ph_pp[f] = rhs_c;
        }
        // forcing row r = m-1 and forward elimination
        if (m >= 2) {
            int r = m - 1;
            float wr = force_row(r, rhs_a, rhs_b, rhs_c, pho_a, pho_b, pho_c,
                                 wold_b, t2_a, t2_b, c2a_a, c2a_b, alt_a,
                                 alt_b, chm_a, chm_b);
            float wf = (wr - d.a[r * st + c] * wf_prev) * d.alpha[r * st + c];
            d.w_pp[r * st + c] = wf;
            wf_prev = wf;
        }
        // shift windows
        rhs_a = rhs_b; rhs_b = rhs_c;
        pho_a = pho_b; pho_b = pho_c;
        t2_a = t2_b; t2_b = t2;
        c2a_a = c2a_b; c2a_b = c2a;
        alt_a = alt_b; alt_b = alt;
        chm_a = chm_b; chm_b = chm;
        wold_b = wold_c;
        ww_k = ww_kp1;
        wdtn_k = wdtn_kp1;
        th_m = th_up;
        ph_m = ph_up;
        wd_prev = wd;
    }
    // top full level nz: rhs (no Omega term), then rows nz-1 and nz
    {
        int f = nz * st + c;
        float pho_c = d.ph_pp[f];
        float wold_c = d.w_pp[f];
        float r = dtau * (d.rph_t[f] + ghalf_eps * wold_c);
        float cfm = d.c1f[nz] * mut + d.c2f[nz];
        float rhs_c = pho_c + r / cfm;
        if (s.top_lid) rhs_c = 0.0f;
        d.ph_pp[f] = rhs_c;
        if (nz >= 2) {
            int rr = nz - 1;
            float wr = force_row(rr, rhs_a, rhs_b, rhs_c, pho_a, pho_b, pho_c,
                                 wold_b, t2_a, t2_b, c2a_a, c2a_b, alt_a,
                                 alt_b, chm_a, chm_b);
            float wf = (wr - d.a[rr * st + c] * wf_prev) * d.alpha[rr * st + c];
            d.w_pp[rr * st + c] = wf;
            wf_prev = wf;
        }
        float wtop;
        if (s.top_lid) {
            wtop = 0.0f;
        } else {
            float dph_dn = one_p * (rhs_c - rhs_b) + one_m * (pho_c - pho_b);
            float bterm = dtau * G_F * c2a_b * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                          / chm_b * dph_dn;
            float cterm = dtau * G_F * (2.0f * d.rdnw[nz - 1] * c2a_b * alt_b * t2_b
                                        + d.c1f[nz] * muave);
            wtop = wold_c + ((dtau * d.rw_t[f] - bterm) - cterm);
        }
        float wf = (wtop - d.a[f] * wf_prev) * d.alpha[f];
        d.w_pp[f] = wf;
    }

    // downward sweep: back substitution, damper, phi'' update, EOS
    float htop = 0.0f, hbot = 0.0f;
    if (s.damp) {
        htop = phs(nz) / G_F;
        hbot = htop - s.zdamp;
    }
    const float c_ph = ((0.5f * dtau) * G_F) * one_p;
    float wb_next = 0.0f;         // back-substituted (undamped) w[k+1]
    float phn_up = 0.0f;          // new phi'' at k+1
    for (int k = nz; k >= 0; --k) {
        int f = k * st + c;
        float phn;
        if (k >= 1) {
            float wb = d.w_pp[f];
            if (k < nz) wb = wb - d.gam[f] * wb_next;
            wb_next = wb;
            float wfin = wb;
            if (s.damp) {
                float hk = phs(k) / G_F;
                if (hk >= hbot) {
                    float arg = (1.5707963f * (hk - hbot)) / s.zdamp;
                    float sn = (float)sin((double)arg);
                    float dw = (s.dampmag * sn) * sn;
                    float cfm = d.c1f[k] * mut + d.c2f[k];
                    wfin = (wb - (dw * cfm) * d.w[f]) / (1.0f + dw);
                }
            }
            d.w_pp[f] = wfin;
            float cfs = d.c1f[k] * muts + d.c2f[k];
            phn = d.ph_pp[f] + (c_ph * wfin) / cfs;
            d.ph_pp[f] = phn;
        } else {
            phn = d.ph_pp[f];
        }
        if (k < nz) {
            int h = f;
            float chm2 = d.c1h[k] * muts + d.c2h[k];
            float th_ref = thr(k, c);
            float alt = d.alt[h];
            float al = -((alt * (d.c1h[k] * mupp)) + d.rdnw[k] * (phn_up - phn)) / chm2;
            d.al_pp[h] = al;
            pout[h] = d.c2a[h] * ((alt * (d.th_pp[h] - (d.c1h[k] * mupp) * th_ref))
                                      / (chm2 * th_ref) - al);
        }
        phn_up = phn;
    }
}

// ---------------------------------------------------------------------------
// Small-step finish (+ EOS of the new state) and lower boundary
// ---------------------------------------------------------------------------

__global__ void k_finish_uv(Dom d)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > nx || j > ny || k >= nz) return;
    if (j < ny) {
        int ic = (i >= nx) ? i - nx : i;
        int ib = (ic == 0) ? nx - 1 : ic - 1;
        int cA = j * nx + ic, cB = j * nx + ib;
        float msA = mut_at(d, cA), msB = mut_at(d, cB);
        float mnA = msA + d.mu_pp[cA], mnB = msB + d.mu_pp[cB];
        float msf = 0.5f * (msA + msB), mnf = 0.5f * (mnA + mnB);
        int ix = iU(k, j, i, ny, nx);
        d.u[ix] = ((d.c1h[k] * msf + d.c2h[k]) * d.u[ix] + d.u_pp[ix])
                  / (d.c1h[k] * mnf + d.c2h[k]);
    }
    if (i < nx) {
        int jc = (j >= ny) ? j - ny : j;
        int jb = (jc == 0) ? ny - 1 : jc - 1;
        int cA = jc * nx + i, cB = jb * nx + i;
        float msA = mut_at(d, cA), msB = mut_at(d, cB);
        float mnA = msA + d.mu_pp[cA], mnB = msB + d.mu_pp[cB];
        float msf = 0.5f * (msA + msB), mnf = 0.5f * (mnA + mnB);
        int ix = iV(k, j, i, ny, nx);
        d.v[ix] = ((d.c1h[k] * msf + d.c2h[k]) * d.v[ix] + d.v_pp[ix])
                  / (d.c1h[k] * mnf + d.c2h[k]);
    }
}

__global__ void k_finish_col(Dom d)
{
    const int nz = d.nz, st = d.ny * d.nx;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int b3 = d.base3d;
    float mus = mut_at(d, c);
    float mupp = d.mu_pp[c];
    float mun = mus + mupp;
    float mup_new = d.mup[c] + mupp;
    float mu_eos = d.mub2d[c] + mup_new;
    float php_lo = 0.0f, thp_lo = 0.0f;
    for (int k = 0; k <= nz; ++k) {
        int f = k * st + c;
        d.w[f] = ((d.c1f[k] * mus + d.c2f[k]) * d.w[f] + d.w_pp[f])
                 / (d.c1f[k] * mun + d.c2f[k]);
        float php_new = d.php[f] + d.ph_pp[f];
        d.php[f] = php_new;
        if (k >= 1) {                            // EOS of level k-1
            int h = (k - 1) * st + c;
            float th = basev(d.thb, k - 1, c, st, b3) + thp_lo;
            float dph = (basev(d.phb, k, c, st, b3) + php_new)
                        - (basev(d.phb, k - 1, c, st, b3) + php_lo);
            float alt = (-dph * d.rdnw[k - 1]) / (d.c1h[k - 1] * mu_eos + d.c2h[k - 1]);
            d.alt[h] = alt;
            d.al[h] = alt - basev(d.alb, k - 1, c, st, b3);
            d.p[h] = P0_F * pow_gamma((RD_F * th) / (P0_F * alt));
        }
        if (k < nz) {
            float thb = basev(d.thb, k, c, st, b3);
            float num = (d.c1h[k] * mus + d.c2h[k]) * (thb + d.thp[f]) + d.th_pp[f];
            float thn = num / (d.c1h[k] * mun + d.c2h[k]) - thb;
            d.thp[f] = thn;
            thp_lo = thn;
        }
        php_lo = php_new;
    }
    d.mup[c] = mup_new;
}

__global__ void k_set_w_surface(Dom d, Sc s)
{
    const int ny = d.ny, nx = d.nx;
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ny * nx) return;
    int j = c / nx, i = c - j * nx;
    int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    auto uc = [&](int f) {
        return (d.cf1 * d.u[iU(0, j, f, ny, nx)] + d.cf2 * d.u[iU(1, j, f, ny, nx)])
               + d.cf3 * d.u[iU(2, j, f, ny, nx)];
    };
    auto vc = [&](int g) {
        return (d.cf1 * d.v[iV(0, g, i, ny, nx)] + d.cf2 * d.v[iV(1, g, i, ny, nx)])
               + d.cf3 * d.v[iV(2, g, i, ny, nx)];
    };
    float h = d.ht[c];
    float dyn = d.ht[jp * nx + i] - h, dys = h - d.ht[jm * nx + i];
    float dxe = d.ht[j * nx + ip] - h, dxw = h - d.ht[j * nx + im];
    d.w[c] = s.hrdy * (dyn * vc(j + 1) + dys * vc(j))
             + s.hrdx * (dxe * uc(i + 1) + dxw * uc(i));
}

__global__ void k_faces(Dom d, float* ru, float* rv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > d.nx || j > d.ny || k >= d.nz) return;
    if (j < d.ny) ru[iU(k, j, i, d.ny, d.nx)] = ru_at(d, k, j, i);
    if (i < d.nx) rv[iV(k, j, i, d.ny, d.nx)] = rv_at(d, k, j, i);
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned absbits(float x)
{
    return __float_as_uint(x) & 0x7fffffffu;
}

__global__ void k_zero_u32(unsigned* p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

// max |x| as ordered uint bits (NaN has the largest pattern, so a NaN
// anywhere survives the maximum).
__global__ void k_health(Dom d, unsigned* out)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    __shared__ unsigned red[4][256];
    unsigned mu = 0, mw = 0, mt = 0, mc = 0;
    int nu = nz * ny * (nx + 1), nw = (nz + 1) * st, nt = nz * st;
    int stride = gridDim.x * blockDim.x;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nw; t += stride) {
        if (t < nu) mu = max(mu, absbits(d.u[t]));
        mw = max(mw, absbits(d.w[t]));
        if (t < nt) {
            mt = max(mt, absbits(d.thp[t]));
            int k = t / st, c = t - k * st;
            float dz = ((basev(d.phb, k + 1, c, st, d.base3d) + d.php[t + st])
                        - (basev(d.phb, k, c, st, d.base3d) + d.php[t])) / G_F;
            mc = max(mc, absbits(d.w[t + st] / dz));
        }
    }
    red[0][threadIdx.x] = mu;
    red[1][threadIdx.x] = mw;
    red[2][threadIdx.x] = mt;
    red[3][threadIdx.x] = mc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 4; ++q)
                red[q][threadIdx.x] = max(red[q][threadIdx.x], red[q][threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x < 4) atomicMax(&out[threadIdx.x], red[threadIdx.x][0]);
}

// FP64 column-mass sum with a fixed summation order (one block).
__global__ void k_mass(Dom d, double* out)
{
    __shared__ double red[1024];
    int st = d.ny * d.nx;
    double acc = 0.0;
    for (int c = threadIdx.x; c < st; c += blockDim.x)
        acc += (double)(d.mub2d[c] + d.mup[c]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

// Optional per-kernel timing: when active, an event is recorded on the
// launch stream after every kernel of arw_step, and the interval since the
// previous event is attributed to that kernel.  Events come from a pool
// created by arw_profile_begin, so the timed step allocates nothing.
enum KernelId {
    KID_STAGE_COL = 0, KID_SLOW_TEND, KID_ADVANCE_UV, KID_ACOUSTIC_COL,
    KID_FINISH_UV, KID_FINISH_COL, KID_SET_W_SURFACE, KID_COPY, KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {
    "k_stage_col", "k_slow_tend", "k_advance_uv", "k_acoustic_col",
    "k_finish_uv", "k_finish_col", "k_set_w_surface", "time_t_copies"};

struct Profiler {
    bool active = false;
    int cap = 0, n = 0;
    hipEvent_t* ev = nullptr;
    int* kid = nullptr;
    bool overflow = false;
} g_prof;

void prof_mark(int id, hipStream_t st)
{
    if (!g_prof.active) return;
    if (g_prof.n >= g_prof.cap) {
        g_prof.overflow = true;
        return;
    }
    g_prof.kid[g_prof.n] = id;
    hipEventRecord(g_prof.ev[g_prof.n], st);
    ++g_prof.n;
}

int32_t check_domain(const Dom* d)
{
    if (!d) return fail(ARW_E_ARG, "null domain");
    if (d->nz < 3 || d->ny < 1 || d->nx < 1)
        return fail(ARW_E_ARG, "need nz >= 3, ny >= 1, nx >= 1");
    long long big = (long long)(d->nz + 1) * (d->ny + 1) * (d->nx + 1);
    if (big >= (1LL << 31))
        return fail(ARW_E_ARG, "domain exceeds 2^31 points per field");
    const void* req[] = {d->u, d->v, d->w, d->thp, d->php, d->mup, d->p,
                         d->al, d->alt, d->u0, d->v0, d->w0, d->thp0,
                         d->php0, d->mup0, d->ru_t, d->rv_t, d->rw_t,
                         d->rth_t, d->rph_t, d->rmu_t, d->u_pp, d->v_pp,
                         d->w_pp, d->th_pp, d->ph_pp, d->mu_pp, d->p_pp,
                         d->p_pp_old, d->ww_pp, d->al_pp, d->thb, d->pb,
                         d->alb, d->phb, d->mub2d, d->ht, d->c1h, d->c2h,
                         d->c1f, d->c2f, d->dnw, d->rdnw, d->dn, d->rdn,
                         d->fnm, d->fnp, d->ww, d->c2a, d->a, d->alpha,
                         d->gam, d->dmdt_rf};
    for (const void* p : req)
        if (!p) return fail(ARW_E_ARG, "null array in arw_domain");
    return ARW_OK;
}

int32_t check_params(const arw_params* p, const Dom* d)
{
    if (!p) return fail(ARW_E_ARG, "null params");
    if (!(p->dx > 0.0) || !(p->dy > 0.0) || !(p->dt > 0.0))
        return fail(ARW_E_ARG, "dx, dy, dt must be positive");
    if (p->time_step_sound < 1 || p->time_step_sound % 2)
        return fail(ARW_E_ARG, "time_step_sound must be even and positive");
    if ((p->khdif > 0.0 || p->kvdif > 0.0)
        && (!d->rdzf_h || !d->rdzc_h || !d->rdzf_w || !d->rdzc_w))
        return fail(ARW_E_ARG, "diffusion spacings required when khdif/kvdif > 0");
    return ARW_OK;
}

inline int32_t launched(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
        return ARW_E_LAUNCH;
    }
    return ARW_OK;
}

inline dim3 col_grid(const Dom& d) { return dim3((d.ny * d.nx + 255) / 256); }
inline dim3 union_grid(const Dom& d, int nlev)
{
    return dim3((d.nx + 1 + 63) / 64, (d.ny + 1 + 3) / 4, nlev);
}

int32_t stage_prepare(const Dom& d, const Sc& s, int do_eos, hipStream_t st)
{
    hipLaunchKernelGGL(k_stage_col, col_grid(d), dim3(256), 0, st, d, s,
                       do_eos, 1, 1);
    int32_t rc = launched("k_stage_col");
    if (rc) return rc;
    prof_mark(KID_STAGE_COL, st);
    hipLaunchKernelGGL(k_slow_tend, union_grid(d, d.nz + 1), dim3(64, 4), 0,
                       st, d, s, 1);
    rc = launched("k_slow_tend");
    prof_mark(KID_SLOW_TEND, st);
    return rc;
}

int32_t substep(const Dom& d, const Sc& s, float* pout, hipStream_t st)
{
    hipLaunchKernelGGL(k_advance_uv, union_grid(d, d.nz), dim3(64, 4), 0, st,
                       d, s);
    int32_t rc = launched("k_advance_uv");
    if (rc) return rc;
    prof_mark(KID_ADVANCE_UV, st);
    hipLaunchKernelGGL(k_acoustic_col, col_grid(d), dim3(256), 0, st, d, s,
                       pout);
    rc = launched("k_acoustic_col");
    prof_mark(KID_ACOUSTIC_COL, st);
    return rc;
}

int32_t finish(const Dom& d, hipStream_t st)
{
    hipLaunchKernelGGL(k_finish_uv, union_grid(d, d.nz), dim3(64, 4), 0, st,
                       d);
    int32_t rc = launched("k_finish_uv");
    if (rc) return rc;
    prof_mark(KID_FINISH_UV, st);
    hipLaunchKernelGGL(k_finish_col, col_grid(d), dim3(256), 0, st, d);
    rc = launched("k_finish_col");
    prof_mark(KID_FINISH_COL, st);
    return rc;
}

}  // namespace

extern "C" {

int32_t arw_abi_version(void) { return ARW_ABI_VERSION; }

const char* arw_last_error(void) { return g_err; }

int32_t arw_calc_p_alpha(const arw_domain* d, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eos, col_grid(*d), dim3(256), 0,
                       (hipStream_t)stream, *d);
    return launched("k_eos");
}

int32_t arw_stage_fluxes(const arw_domain* d, const arw_params* prm,
                         float* ru, float* rv, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !ru || !rv) return fail(ARW_E_ARG, "null argument");
    Sc s = make_sc(prm, prm->dt, 1);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_stage_col, col_grid(*d), dim3(256), 0, st, *d, s,
                       0, 0, 0);
    if ((rc = launched("k_stage_col"))) return rc;
    // materialise ru/rv with the same device functions the kernels use
    hipLaunchKernelGGL(k_faces, union_grid(*d, d->nz), dim3(64, 4), 0, st,
                       *d, ru, rv);
    return launched("k_faces");
}

static int32_t flux_div(int which, const float* q, const float* ru,
                        const float* rv, const float* rw, float* tend,
                        const float* rdz, const float* fnm, const float* fnp,
                        float dx_inv, float dy_inv, int32_t nz, int32_t ny,
                        int32_t nx, void* stream)
{
    if (!q || !ru || !rv || !rw || !tend || !rdz || !fnm || !fnp)
        return fail(ARW_E_ARG, "null argument");
    if (nz < 1 || ny < 1 || nx < 1) return fail(ARW_E_ARG, "bad extents");
    dim3 grid((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz + 1);
    hipLaunchKernelGGL(k_flux_div, grid, dim3(64, 4), 0, (hipStream_t)stream,
                       which, q, ru, rv, rw, tend, rdz, fnm, fnp, dx_inv,
                       dy_inv, nz, ny, nx);
    return launched("k_flux_div");
}

int32_t arw_flux_div_scalar(const float* q, const float* ru, const float* rv,
                            const float* rw, float* tend, const float* rdnw,
                            const float* fnm, const float* fnp, float dx_inv,
                            float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                            void* stream)
{
    return flux_div(0, q, ru, rv, rw, tend, rdnw, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t arw_flux_div_u(const float* u, const float* ru, const float* rv,
                       const float* rw, float* tend, const float* rdnw,
                       const float* fnm, const float* fnp, float dx_inv,
                       float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void* stream)
{
    return flux_div(1, u, ru, rv, rw, tend, rdnw, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t arw_flux_div_v(const float* v, const float* ru, const float* rv,
                       const float* rw, float* tend, const float* rdnw,
                       const float* fnm, const float* fnp, float dx_inv,
                       float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void* stream)
{
    return flux_div(2, v, ru, rv, rw, tend, rdnw, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t arw_flux_div_w(const float* w, const float* ru, const float* rv,
                       const float* rw, float* tend, const float* rdn,
                       const float* fnm, const float* fnp, float dx_inv,
                       float dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void* stream)
{
    return flux_div(3, w, ru, rv, rw, tend, rdn, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t arw_stage_prepare(const arw_domain* d, const arw_params* prm,
                          double dtau, int32_t do_eos, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if ((rc = check_params(prm, d))) return rc;
    Sc s = make_sc(prm, dtau, 1);
    return stage_prepare(*d, s, do_eos, (hipStream_t)stream);
}

int32_t arw_calc_coefs(const arw_domain* d, const arw_params* prm,
                       double dtau, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm) return fail(ARW_E_ARG, "null params");
    Sc s = make_sc(prm, dtau, 1);
    hipLaunchKernelGGL(k_stage_col, col_grid(*d), dim3(256), 0,
                       (hipStream_t)stream, *d, s, 0, 0, 1);
    return launched("k_stage_col");
}

int32_t arw_acoustic_substep(const arw_domain* d, const arw_params* prm,
                             double dtau, int32_t first, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm) return fail(ARW_E_ARG, "null params");
    Sc s = make_sc(prm, dtau, first);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_advance_uv, union_grid(*d, d->nz), dim3(64, 4), 0,
                       st, *d, s);
    if ((rc = launched("k_advance_uv"))) return rc;
    // the reference keeps p_pp_old = p'' before the column update
    size_t bytes = (size_t)d->nz * d->ny * d->nx * sizeof(float);
    if (hipMemcpyAsync(d->p_pp_old, d->p_pp, bytes, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return fail(ARW_E_LAUNCH, "p_pp_old copy failed");
    hipLaunchKernelGGL(k_acoustic_col, col_grid(*d), dim3(256), 0, st, *d, s,
                       d->p_pp);
    return launched("k_acoustic_col");
}

int32_t arw_small_step_finish(const arw_domain* d, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    return finish(*d, (hipStream_t)stream);
}

int32_t arw_set_w_surface(const arw_domain* d, const arw_params* prm,
                          void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm) return fail(ARW_E_ARG, "null params");
    Sc s = make_sc(prm, prm->dt, 1);
    hipLaunchKernelGGL(k_set_w_surface, col_grid(*d), dim3(256), 0,
                       (hipStream_t)stream, *d, s);
    return launched("k_set_w_surface");
}

int32_t arw_step(const arw_domain* d, const arw_params* prm, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if ((rc = check_params(prm, d))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t nz = d->nz, ny = d->ny, nx = d->nx;
    struct { float* dst; const float* src; size_t n; } copies[] = {
        {d->u0, d->u, nz * ny * (nx + 1)}, {d->v0, d->v, nz * (ny + 1) * nx},
        {d->w0, d->w, (nz + 1) * ny * nx}, {d->thp0, d->thp, nz * ny * nx},
        {d->php0, d->php, (nz + 1) * ny * nx}, {d->mup0, d->mup, ny * nx}};
    prof_mark(KID_COUNT, st);                    // opening event
    for (auto& cp : copies)
        if (hipMemcpyAsync(cp.dst, cp.src, cp.n * sizeof(float),
                           hipMemcpyDeviceToDevice, st) != hipSuccess)
            return fail(ARW_E_LAUNCH, "time-t copy failed");
    prof_mark(KID_COPY, st);
    const int ns = prm->time_step_sound;
    const int nsub[3] = {1, ns / 2 > 1 ? ns / 2 : 1, ns};
    const double dtaus[3] = {prm->dt / 3.0, prm->dt / ns, prm->dt / ns};
    for (int stage = 0; stage < 3; ++stage) {
        Sc s = make_sc(prm, dtaus[stage], 1);
        if ((rc = stage_prepare(*d, s, stage == 0, st))) return rc;
        Dom dd = *d;
        for (int n = 0; n < nsub[stage]; ++n) {
            Sc sn = make_sc(prm, dtaus[stage], n == 0);
            // new p'' goes to the buffer holding the dead p''(n-1); the two
            // names swap afterwards
            if ((rc = substep(dd, sn, dd.p_pp_old, st))) return rc;
            float* t = dd.p_pp;
            dd.p_pp = dd.p_pp_old;
            dd.p_pp_old = t;
        }
        if ((rc = finish(*d, st))) return rc;
    }
    Sc s = make_sc(prm, prm->dt, 1);
    hipLaunchKernelGGL(k_set_w_surface, col_grid(*d), dim3(256), 0, st, *d,
                       s);
    rc = launched("k_set_w_surface");
    prof_mark(KID_SET_W_SURFACE, st);
    return rc;
}

int32_t arw_profile_begin(int32_t max_events)
{
    if (g_prof.active) return fail(ARW_E_ARG, "profiling already active");
    if (max_events < 2) return fail(ARW_E_ARG, "max_events must be >= 2");
    g_prof.ev = new (std::nothrow) hipEvent_t[max_events];
    g_prof.kid = new (std::nothrow) int[max_events];
    if (!g_prof.ev || !g_prof.kid) return fail(ARW_E_ARG, "out of memory");
    for (int i = 0; i < max_events; ++i)
        if (hipEventCreate(&g_prof.ev[i]) != hipSuccess) {
            for (int j = 0; j < i; ++j) hipEventDestroy(g_prof.ev[j]);
            delete[] g_prof.ev;
            delete[] g_prof.kid;
            g_prof = Profiler();
            return fail(ARW_E_LAUNCH, "hipEventCreate failed");
        }
    g_prof.cap = max_events;
    g_prof.n = 0;
    g_prof.overflow = false;
    g_prof.active = true;
    return ARW_OK;
}

int32_t arw_profile_end(double* ms, int32_t* launches, int32_t n_kernels)
{
    if (!g_prof.active) return fail(ARW_E_ARG, "profiling not active");
    int32_t rc = ARW_OK;
    for (int k = 0; k < n_kernels; ++k) {
        if (ms) ms[k] = 0.0;
        if (launches) launches[k] = 0;
    }
    if (g_prof.n > 0 && hipEventSynchronize(g_prof.ev[g_prof.n - 1]) != hipSuccess)
        rc = fail(ARW_E_LAUNCH, "hipEventSynchronize failed");
    for (int i = 1; rc == ARW_OK && i < g_prof.n; ++i) {
        int id = g_prof.kid[i];
        if (g_prof.kid[i - 1] == KID_COUNT && id == KID_COUNT) continue;
        if (id >= KID_COUNT || id >= n_kernels) continue;
        float t = 0.0f;
        hipEventElapsedTime(&t, g_prof.ev[i - 1], g_prof.ev[i]);
        if (ms) ms[id] += t;
        if (launches) launches[id] += 1;
    }
    if (g_prof.overflow && rc == ARW_OK)
        rc = fail(ARW_E_ARG, "profile event pool overflowed");
    for (int i = 0; i < g_prof.cap; ++i) hipEventDestroy(g_prof.ev[i]);
    delete[] g_prof.ev;
    delete[] g_prof.kid;
    g_prof = Profiler();
    return rc;
}

const char* arw_kernel_name(int32_t id)
{
    return (id >= 0 && id < KID_COUNT) ? kKernelNames[id] : "";
}

int32_t arw_kernel_count(void) { return KID_COUNT; }

int32_t arw_health(const arw_domain* d, float* out, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!out) return fail(ARW_E_ARG, "null output");
    hipStream_t st = (hipStream_t)stream;
    unsigned* o = reinterpret_cast<unsigned*>(out);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, o, 4);
    if ((rc = launched("k_zero_u32"))) return rc;
    int n = (d->nz + 1) * d->ny * d->nx;
    int blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_health, dim3(blocks), dim3(256), 0, st, *d, o);
    return launched("k_health");
}

int32_t arw_domain_mass(const arw_domain* d, double* out, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!out) return fail(ARW_E_ARG, "null output");
    hipLaunchKernelGGL(k_mass, dim3(1), dim3(1024), 0, (hipStream_t)stream,
                       *d, out);
    return launched("k_mass");
}

}  // extern "C"
