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
#include "march.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

namespace ARW_NS {

static thread_local char g_err[512] = "";

static int32_t fail(int32_t code, const char* msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return code;
}

// ---------------------------------------------------------------------------
// Shared device pieces
// ---------------------------------------------------------------------------

__device__ __forceinline__ real mut_at(const Dom& d, int c)
{
    return d.mub2d[c] + d.mup[c];
}

// Coupled x mass flux at u face f (0..nx) of row j, level k:
// (c1h*<mu>_x + c2h)*u with <mu>_x = 0.5*(mu[f] + mu[f-1]) (periodic).
__device__ __forceinline__ real ru_at(const Dom& d, int k, int j, int f)
{
    const int nx = d.nx, ny = d.ny;
    int a = (f >= nx) ? f - nx : f;
    int b = (f == 0) ? nx - 1 : f - 1;
    real mux = R(0.5) * (mut_at(d, j * nx + a) + mut_at(d, j * nx + b));
    return (d.c1h[k] * mux + d.c2h[k]) * d.u[iU(k, j, f, ny, nx)];
}

__device__ __forceinline__ real rv_at(const Dom& d, int k, int g, int i)
{
    const int nx = d.nx, ny = d.ny;
    int a = (g >= ny) ? g - ny : g;
    int b = (g == 0) ? ny - 1 : g - 1;
    real muy = R(0.5) * (mut_at(d, a * nx + i) + mut_at(d, b * nx + i));
    return (d.c1h[k] * muy + d.c2h[k]) * d.v[iV(k, g, i, ny, nx)];
}

// Vertical face flux of a half-level field at w-level kf (reference
// advection.cu:189-201): zero at the boundaries, fnm/fnp 2nd order one
// face in, Omega-signed flux3 inside.
template <class Q>
__device__ __forceinline__ real zface_half(Q q, real vel, int kf, int nz,
                                            const real* __restrict__ fnm,
                                            const real* __restrict__ fnp)
{
    if (kf == 0 || kf == nz) return R(0.0);
    if (kf == 1 || kf == nz - 1)
        return vel * (fnm[kf] * q(kf) + fnp[kf] * q(kf - 1));
    return flux3(q(kf - 2), q(kf - 1), q(kf), q(kf + 1), vel);
}

__device__ __forceinline__ real div3(real dfx, real dfy, real dfz,
                                      real rdx, real rdy, real rdz)
{
    real t = -dfx * rdx;
    t = t - dfy * rdy;
    return t - dfz * rdz;
}

// ---------------------------------------------------------------------------
// Advection device functions (periodic, msf = 1)
// ---------------------------------------------------------------------------

// Scalar q at mass point (k,j,i); q(kk,jj,ii) accessor with wrapped
// indices; RU(k,j,f), RV(k,g,i), RW(kf,j,i) advecting fluxes.
template <class Q, class RU, class RV, class RW>
__device__ real adv_scalar(Q q, RU ru, RV rv, RW rw, int k, int j, int i,
                            int nz, int ny, int nx,
                            const real* __restrict__ rdnw,
                            const real* __restrict__ fnm,
                            const real* __restrict__ fnp,
                            real rdx, real rdy)
{
    real x[7], y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        x[o] = q(k, j, pmod(i + o - 3, nx));
        y[o] = q(k, pmod(j + o - 3, ny), i);
    }
    real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], ru(k, j, i));
    real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], ru(k, j, i + 1));
    real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], rv(k, j, i));
    real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], rv(k, j + 1, i));
    auto qz = [&](int kk) { return q(kk, j, i); };
    real fz0 = zface_half(qz, rw(k, j, i), k, nz, fnm, fnp);
    real fz1 = zface_half(qz, rw(k + 1, j, i), k + 1, nz, fnm, fnp);
    return div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, rdx, rdy, rdnw[k]);
}

// u momentum at u point (k, j, i), i = 0..nx (reference advection.cu:
// 572-611).  u(kk,jj,ii) reads the u array at a wrapped core column.
template <class U, class RU, class RV, class RW>
__device__ real adv_u(U u, RU ru, RV rv, RW rw, int k, int j, int i,
                       int nz, int ny, int nx,
                       const real* __restrict__ rdnw,
                       const real* __restrict__ fnm,
                       const real* __restrict__ fnp, real rdx, real rdy)
{
    int ic = (i >= nx) ? i - nx : i;
    int im = (ic == 0) ? nx - 1 : ic - 1;
    // x faces at mass centres im (left) and ic (right)
    real x[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) x[o] = u(k, j, pmod(ic + o - 3, nx));
    real vl = R(0.5) * (ru(k, j, im) + ru(k, j, im + 1));
    real vr = R(0.5) * (ru(k, j, ic) + ru(k, j, ic + 1));
    real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vl);
    real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], vr);
    real y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) y[o] = u(k, pmod(j + o - 3, ny), ic);
    real v0 = R(0.5) * (rv(k, j, im) + rv(k, j, ic));
    real v1 = R(0.5) * (rv(k, j + 1, im) + rv(k, j + 1, ic));
    real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
    real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
    auto qz = [&](int kk) { return u(kk, j, ic); };
    real w0 = R(0.5) * (rw(k, j, im) + rw(k, j, ic));
    real w1 = R(0.5) * (rw(k + 1, j, im) + rw(k + 1, j, ic));
    real fz0 = zface_half(qz, w0, k, nz, fnm, fnp);
    real fz1 = zface_half(qz, w1, k + 1, nz, fnm, fnp);
    return div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, rdx, rdy, rdnw[k]);
}

// v momentum at v point (k, j, i), j = 0..ny (reference advection.cu:
// 754-794).
template <class V, class RU, class RV, class RW>
__device__ real adv_v(V v, RU ru, RV rv, RW rw, int k, int j, int i,
                       int nz, int ny, int nx,
                       const real* __restrict__ rdnw,
                       const real* __restrict__ fnm,
                       const real* __restrict__ fnp, real rdx, real rdy)
{
    int jc = (j >= ny) ? j - ny : j;
    int jm = (jc == 0) ? ny - 1 : jc - 1;
    real x[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) x[o] = v(k, jc, pmod(i + o - 3, nx));
    real u0 = R(0.5) * (ru(k, jm, i) + ru(k, jc, i));
    real u1 = R(0.5) * (ru(k, jm, i + 1) + ru(k, jc, i + 1));
    real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0);
    real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1);
    real y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) y[o] = v(k, pmod(jc + o - 3, ny), i);
    real vl = R(0.5) * (rv(k, jm, i) + rv(k, jm + 1, i));
    real vr = R(0.5) * (rv(k, jc, i) + rv(k, jc + 1, i));
    real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], vl);
    real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], vr);
    auto qz = [&](int kk) { return v(kk, jc, i); };
    real w0 = R(0.5) * (rw(k, jm, i) + rw(k, jc, i));
    real w1 = R(0.5) * (rw(k + 1, jm, i) + rw(k + 1, jc, i));
    real fz0 = zface_half(qz, w0, k, nz, fnm, fnp);
    real fz1 = zface_half(qz, w1, k + 1, nz, fnm, fnp);
    return div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, rdx, rdy, rdnw[k]);
}

// w momentum at interior w level k = 1..nz-1 (reference advection.cu:
// 1016-1050).
template <class W, class RU, class RV, class RW>
__device__ real adv_w(W w, RU ru, RV rv, RW rw, int k, int j, int i,
                       int nz, int ny, int nx,
                       const real* __restrict__ rdn,
                       const real* __restrict__ fnm,
                       const real* __restrict__ fnp, real rdx, real rdy)
{
    real x[7], y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        x[o] = w(k, j, pmod(i + o - 3, nx));
        y[o] = w(k, pmod(j + o - 3, ny), i);
    }
    real u0 = fnm[k] * ru(k, j, i) + fnp[k] * ru(k - 1, j, i);
    real u1 = fnm[k] * ru(k, j, i + 1) + fnp[k] * ru(k - 1, j, i + 1);
    real v0 = fnm[k] * rv(k, j, i) + fnp[k] * rv(k - 1, j, i);
    real v1 = fnm[k] * rv(k, j + 1, i) + fnp[k] * rv(k - 1, j + 1, i);
    real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0);
    real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1);
    real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
    real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
    real fz[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        int m = k - 1 + s;
        real vel = R(0.5) * (rw(m, j, i) + rw(m + 1, j, i));
        if (m == 0 || m == nz - 1)
            fz[s] = R(0.5) * vel * (w(m + 1, j, i) + w(m, j, i));
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
    const real* p;
    int ny, nx;      // row count and row width of the array
    __device__ real operator()(int k, int j, int i) const
    {
        return p[(k * ny + j) * nx + i];
    }
};

template <int WHICH>
__global__ void k_flux_div(const real* __restrict__ q,
                           const real* __restrict__ ru,
                           const real* __restrict__ rv,
                           const real* __restrict__ rw,
                           real* __restrict__ tend,
                           const real* __restrict__ rdz,
                           const real* __restrict__ fnm,
                           const real* __restrict__ fnp, real rdx,
                           real rdy, int nz, int ny, int nx)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    Arr3 RU{ru, ny, nx + 1}, RV{rv, ny + 1, nx}, RW{rw, ny, nx};
    if (WHICH == 0) {
        if (i >= nx || j >= ny || k >= nz) return;
        Arr3 Q{q, ny, nx};
        tend[iM(k, j, i, ny, nx)] += adv_scalar(Q, RU, RV, RW, k, j, i, nz,
                                                ny, nx, rdz, fnm, fnp, rdx,
                                                rdy);
    } else if (WHICH == 1) {
        if (i > nx || j >= ny || k >= nz) return;
        Arr3 Q{q, ny, nx + 1};
        tend[iU(k, j, i, ny, nx)] += adv_u(Q, RU, RV, RW, k, j, i, nz, ny,
                                           nx, rdz, fnm, fnp, rdx, rdy);
    } else if (WHICH == 2) {
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
// Stage column kernel: EOS, Omega, small-step init, coefficients
// ---------------------------------------------------------------------------

__device__ __forceinline__ void eos_level(const Dom& d, int k, int c, int st,
                                          real mu)
{
    real th = basev(d.thb, k, c, st, d.base3d) + d.thp[k * st + c];
    if (d.qv) th = th * (R(1.0) + RVOVRD_F * d.qv[k * st + c]);
    real ph_hi = basev(d.phb, k + 1, c, st, d.base3d)
                  + d.php[(k + 1) * st + c];
    real ph_lo = basev(d.phb, k, c, st, d.base3d) + d.php[k * st + c];
    real dph = ph_hi - ph_lo;
    real alt, al;
    eos_alpha(d, k, mu, dph, basev(d.alb, k, c, st, d.base3d), d.c1h[k],
              d.c2h[k], d.rdnw[k], alt, al);
    d.alt[k * st + c] = alt;
    d.al[k * st + c] = al;
    d.p[k * st + c] = P0_F * pow_gamma((RD_F * th) / (P0_F * alt));
}

// The precision form's EOS of one cell (eos_pert), p' kept in d.pp.
__device__ __forceinline__ void eos_level_pert(const Dom& d, int k, int c,
                                               int st, real mu)
{
    const int x = k * st + c;
    const EosPert e = eos_pert(
        d.thp[x], basev(d.thb, k, c, st, d.base3d), d.qv ? d.qv[x] : R(0.0),
        d.qv != nullptr, basev(d.alb, k, c, st, d.base3d),
        basev(d.pb, k, c, st, d.base3d), d.mup[c], d.c1h[k] * mu + d.c2h[k],
        d.c1h[k], d.rdnw[k], d.php[x + st] - d.php[x]);
    d.alt[x] = e.alt;
    d.al[x] = e.al;
    d.p[x] = e.p;
    d.pp[x] = e.pp;
}

template <bool PREC>
__global__ void k_eos(Dom d)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int st = d.ny * d.nx;
    if (c >= st) return;
    real mu = mut_at(d, c);
    for (int k = 0; k < d.nz; ++k) {
        if (PREC) eos_level_pert(d, k, c, st, mu);
        else eos_level(d, k, c, st, mu);
    }
}

template <class SRC>
__global__ void k_set_w_surface(typename SRC::Arg a, Sc s)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const int ny = d.ny, nx = d.nx;
    int c = g * blockDim.x + threadIdx.x;
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
    real h = d.ht[c];
    real dyn = d.ht[jp * nx + i] - h, dys = h - d.ht[jm * nx + i];
    real dxe = d.ht[j * nx + ip] - h, dxw = h - d.ht[j * nx + im];
    d.w[c] = s.hrdy * (dyn * vc(j + 1) + dys * vc(j))
             + s.hrdx * (dxe * uc(i + 1) + dxw * uc(i));
}

__global__ void k_faces(Dom d, real* ru, real* rv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > d.nx || j > d.ny || k >= d.nz) return;
    if (j < d.ny) ru[iU(k, j, i, d.ny, d.nx)] = ru_at(d, k, j, i);
    if (i < d.nx) rv[iV(k, j, i, d.ny, d.nx)] = rv_at(d, k, j, i);
}

hipError_t run_eos(const Dom& d, hipStream_t st)
{
    hipLaunchKernelGGL(k_eos<false>, dim3((d.ny * d.nx + 255) / 256),
                       dim3(256), 0, st, d);
    return hipGetLastError();
}

hipError_t run_faces(const Dom& d, real* ru, real* rv, hipStream_t st)
{
    dim3 grid((d.nx + 1 + 63) / 64, (d.ny + 1 + 3) / 4, d.nz);
    hipLaunchKernelGGL(k_faces, grid, dim3(64, 4), 0, st, d, ru, rv);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Moist scalar transport (every species), periodic, msf = 1
// ---------------------------------------------------------------------------

// WRF sumflux (reference acoustic.cu:1437-1467), the per-substep sums of
// u'', v'', Omega'', runs inside the substep pass (k_uv, k_column with
// Sc::sum set).
// WRF sumflux at the last substep (reference acoustic.cu:1469-1489): the
// substep mean plus the stage-reference coupled fluxes ru, rv (from the
// stage state d.u/d.v/d.mup) and Omega (d.ww).
__global__ void k_sumflux_finish(Dom d, real nsub)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    if (j < ny && k < nz) {
        const int x = iU(k, j, i, ny, nx);
        d.ru_m[x] = d.ru_m[x] / nsub + ru_at(d, k, j, i);
    }
    if (i < nx && k < nz) {
        const int x = iV(k, j, i, ny, nx);
        d.rv_m[x] = d.rv_m[x] / nsub + rv_at(d, k, j, i);
    }
    if (i < nx && j < ny) {
        const int x = iM(k, j, i, ny, nx);
        d.ww_m[x] = d.ww_m[x] / nsub + d.ww[x];
    }
}

// WRF calc_cq at the stage state (reference acoustic.cu:84-158): cq =
// 1 / (1 + 0.5 qtot), with qtot the first n moist mass species (qv, qc,
// qr, then the loading extras) summed over the two cells around each u
// face, v face and interior w level; 1 at the surface and the lid.
__global__ void k_calc_cq(Dom d, int n)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    auto pair = [&](int a, int b) {
        real q = R(0.0);
#pragma unroll
        for (int e = 0; e < 3 + 4; ++e) {
            if (e >= n) break;
            const real* x = e == 0 ? d.qv : (e == 1 ? d.qc : (e == 2 ? d.qr
                                                  : d.qx[e - 3]));
            q = q + x[a];
            q = q + x[b];
        }
        return R(1.0) / (R(1.0) + R(0.5) * q);
    };
    if (j < ny && k < nz) {
        const int a = (i >= nx) ? i - nx : i, b = (a == 0) ? nx - 1 : a - 1;
        d.cqu[iU(k, j, i, ny, nx)] = pair(k * st + j * nx + a,
                                          k * st + j * nx + b);
    }
    if (i < nx && k < nz) {
        const int a = (j >= ny) ? j - ny : j, b = (a == 0) ? ny - 1 : a - 1;
        d.cqv[iV(k, j, i, ny, nx)] = pair(k * st + a * nx + i,
                                          k * st + b * nx + i);
    }
    if (i < nx && j < ny) {
        const int c = j * nx + i;
        d.cqw[k * st + c] = (k == 0 || k == nz)
                                ? R(1.0)
                                : pair(k * st + c, (k - 1) * st + c);
    }
}

// Post-acoustic stage mu' of column c: the value small_step_finish
// stores (mup + mu''), formed here so the scalar update can run before
// the finish overwrites the time-t mass on the last stage.
__device__ __forceinline__ real mup_new_at(const Dom& d, int c)
{
    return d.mup[c] + d.mu_pp[c];
}

// WRF rk_update_scalar (reference moist.py:385-398): (C(mu_t) q_t + dt
// tend) / C(mu_new), clamped at zero on the final stage (NaN kept, -0
// becomes +0).
__device__ __forceinline__ real scalar_update(const Dom& d, int k, int c,
                                               real q0, real tend,
                                               const real* mup0, real dt,
                                               int clamp)
{
    const real mu0 = d.mub2d[c] + mup0[c];
    const real mu = d.mub2d[c] + mup_new_at(d, c);
    const real chm0 = d.c1h[k] * mu0 + d.c2h[k];
    const real chm = d.c1h[k] * mu + d.c2h[k];
    real v = chm0 * q0;
    const real t = dt * tend;
    v = v + t;
    v = v / chm;
    if (clamp && !(v > R(0.0)) && v == v) v = R(0.0);
    return v;
}

// Unlimited stage update of one scalar, one thread per point: tend = div
// of qcur with the averaged fluxes ru_m, rv_m, ww_m; qdst = update from
// q0.  Used where the domain has too few columns for the column marches.
__global__ void k_scalar_update(Dom d, const real* __restrict__ qcur,
                                const real* __restrict__ q0,
                                real* __restrict__ qdst,
                                const real* __restrict__ mup0, real dt,
                                int clamp, real rdx, real rdy,
                                const real* __restrict__ fixed)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    Arr3 Q{qcur, ny, nx}, RU{d.ru_m, ny, nx + 1}, RV{d.rv_m, ny + 1, nx},
        RW{d.ww_m, ny, nx};
    real tend = R(0.0) + adv_scalar(Q, RU, RV, RW, k, j, i, nz, ny, nx,
                                   d.rdnw, d.fnm, d.fnp, rdx, rdy);
    const int x = iM(k, j, i, ny, nx);
    if (fixed) tend = tend + fixed[x];       // held mixing (moist.py:384-386)
    qdst[x] = scalar_update(d, k, j * nx + i, q0[x], tend, mup0, dt, clamp);
}

// The held tendency folded into the time-t scalar ahead of the final
// stage's limiter, in place: h = q0 + dt h / C(mu_t) (reference moist.py:
// 293-346 _pd_fold_sources).
__global__ void k_pd_fold(Dom d, const real* __restrict__ q0,
                          real* __restrict__ h, const real* __restrict__ mup0,
                          real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int c = j * nx + i, x = iM(k, j, i, ny, nx);
    const real chm0 = d.c1h[k] * (d.mub2d[c] + mup0[c]) + d.c2h[k];
    h[x] = q0[x] + (dt * h[x]) / chm0;
}

// WRF flux_upwind (reference pd_advection.cu:101-108).
__device__ __forceinline__ real flux_upwind(real qm1, real q0v, real cr)
{
    return (R(0.5) * fmin(R(1.0), cr + fabs(cr))) * qm1
           + (R(0.5) * fmax(-R(1.0), cr - fabs(cr))) * q0v;
}

// PD face fluxes (reference pd_advection.cu pd_fluxes, periodic, no
// msf): the CFL-clamped upwind flux of q0 and the correction to the
// unlimited flux of qcur, at x faces (k, j, i<=nx), y faces (k, j<=ny, i)
// and eta faces (k<=nz, j, i).  mut is the post-acoustic stage mass.
__global__ void k_pd_fluxes(Dom d, const real* __restrict__ qcur,
                            const real* __restrict__ q0, real dx, real dy,
                            real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    auto mut = [&](int jj, int ii) {
        return d.mub2d[jj * nx + ii] + mup_new_at(d, jj * nx + ii);
    };
    auto q0a = [&](int kk, int jj, int ii) { return q0[iM(kk, jj, ii, ny, nx)]; };
    auto qc = [&](int kk, int jj, int ii) { return qcur[iM(kk, jj, ii, ny, nx)]; };
    if (j < ny && k < nz) {
        const int a = (i >= nx) ? i - nx : i, b = (i == 0) ? nx - 1 : i - 1;
        const real vel = d.ru_m[iU(k, j, i, ny, nx)];
        const real muf = (d.c1h[k] * R(0.5)) * (mut(j, a) + mut(j, b)) + d.c2h[k];
        const real cr = ((vel * dt) / dx) / muf;
        const real fl = (muf * (dx / dt)) * flux_upwind(q0a(k, j, b), q0a(k, j, a), cr);
        real x[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) x[o] = qc(k, j, pmod(i + o - 3, nx));
        const real fh = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vel);
        d.pd_fxl[iU(k, j, i, ny, nx)] = fl;
        d.pd_fxc[iU(k, j, i, ny, nx)] = fh - fl;
    }
    if (i < nx && k < nz) {
        const int a = (j >= ny) ? j - ny : j, b = (j == 0) ? ny - 1 : j - 1;
        const real vel = d.rv_m[iV(k, j, i, ny, nx)];
        const real muf = (d.c1h[k] * R(0.5)) * (mut(a, i) + mut(b, i)) + d.c2h[k];
        const real cr = ((vel * dt) / dy) / muf;
        const real fl = (muf * (dy / dt)) * flux_upwind(q0a(k, b, i), q0a(k, a, i), cr);
        real y[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) y[o] = qc(k, pmod(j + o - 3, ny), i);
        const real fh = flux5(y[0], y[1], y[2], y[3], y[4], y[5], vel);
        d.pd_fyl[iV(k, j, i, ny, nx)] = fl;
        d.pd_fyc[iV(k, j, i, ny, nx)] = fh - fl;
    }
    if (i < nx && j < ny) {
        real fl = R(0.0), fc = R(0.0);
        if (k > 0 && k < nz) {
            const real vel = d.ww_m[iM(k, j, i, ny, nx)];
            const real dz = R(2.0) / (d.rdnw[k] + d.rdnw[k - 1]);
            const real muf = d.c1h[k] * mut(j, i) + d.c2h[k];
            const real cr = (vel * dt) / (dz * muf);
            fl = (muf * (dz / dt)) * flux_upwind(q0a(k - 1, j, i), q0a(k, j, i), cr);
            auto qz = [&](int kk) { return qc(kk, j, i); };
            fc = zface_half(qz, vel, k, nz, d.fnm, d.fnp) - fl;
        }
        d.pd_fzl[iM(k, j, i, ny, nx)] = fl;
        d.pd_fzc[iM(k, j, i, ny, nx)] = fc;
    }
}

// PD renormalisation factor of each cell (reference pd_advection.cu
// pd_scale, periodic no-msf branch).
__global__ void k_pd_scale(Dom d, const real* __restrict__ q0,
                           const real* __restrict__ mup0, real dxi,
                           real dyi, real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int c = j * nx + i;
    const real chm0 = d.c1h[k] * (d.mub2d[c] + mup0[c]) + d.c2h[k];
    const int u0 = iU(k, j, i, ny, nx), v0 = iV(k, j, i, ny, nx);
    const int z0 = iM(k, j, i, ny, nx), st = ny * nx;
    const real dl = (dxi * (d.pd_fxl[u0 + 1] - d.pd_fxl[u0])
                      + dyi * (d.pd_fyl[v0 + nx] - d.pd_fyl[v0]))
                     + d.rdnw[k] * (d.pd_fzl[z0 + st] - d.pd_fzl[z0]);
    const real ph_low = chm0 * q0[z0] - dt * dl;
    const real fo = dt * ((dxi * (fmax(R(0.0), d.pd_fxc[u0 + 1])
                                   - fmin(R(0.0), d.pd_fxc[u0]))
                            + dyi * (fmax(R(0.0), d.pd_fyc[v0 + nx])
                                     - fmin(R(0.0), d.pd_fyc[v0])))
                           + d.rdnw[k] * (fmin(R(0.0), d.pd_fzc[z0 + st])
                                          - fmax(R(0.0), d.pd_fzc[z0])));
    real r = R(1.0);
    if (fo > ph_low) r = fmin(R(1.0), fmax(R(0.0), ph_low / (fo + R(1e-20))));
    d.pd_sc[z0] = r;
}

// Renormalised PD divergence and the coupled update (reference
// pd_advection.cu pd_renorm_apply + moist.py:385-398, final stage, clamp):
// each correction flux takes its donor cell's scale.
__global__ void k_pd_apply(Dom d, const real* __restrict__ q0,
                           real* __restrict__ qdst,
                           const real* __restrict__ mup0, real dxi,
                           real dyi, real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int st = ny * nx;
    const int u0 = iU(k, j, i, ny, nx), v0 = iV(k, j, i, ny, nx);
    const int z0 = iM(k, j, i, ny, nx);
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    const real* sc = d.pd_sc;
    const real s0 = sc[z0];
    const real fxc_l = d.pd_fxc[u0], fxc_r = d.pd_fxc[u0 + 1];
    const real sx_l = (fxc_l > R(0.0)) ? sc[iM(k, j, im, ny, nx)]
                                      : ((fxc_l < R(0.0)) ? s0 : R(1.0));
    const real sx_r = (fxc_r > R(0.0)) ? s0
                       : ((fxc_r < R(0.0)) ? sc[iM(k, j, ip, ny, nx)] : R(1.0));
    const real fyc_l = d.pd_fyc[v0], fyc_r = d.pd_fyc[v0 + nx];
    const real sy_l = (fyc_l > R(0.0)) ? sc[iM(k, jm, i, ny, nx)]
                                      : ((fyc_l < R(0.0)) ? s0 : R(1.0));
    const real sy_r = (fyc_r > R(0.0)) ? s0
                       : ((fyc_r < R(0.0)) ? sc[iM(k, jp, i, ny, nx)] : R(1.0));
    const real fzc_b = d.pd_fzc[z0], fzc_t = d.pd_fzc[z0 + st];
    const real sz_b = (fzc_b > R(0.0)) ? s0
                       : ((fzc_b < R(0.0)) ? sc[z0 - st] : R(1.0));
    const real sz_t = (fzc_t < R(0.0)) ? s0
                       : ((fzc_t > R(0.0)) ? sc[z0 + st] : R(1.0));
    const real gx = (sx_r * fxc_r + d.pd_fxl[u0 + 1]) - (sx_l * fxc_l + d.pd_fxl[u0]);
    const real gy = (sy_r * fyc_r + d.pd_fyl[v0 + nx]) - (sy_l * fyc_l + d.pd_fyl[v0]);
    const real gz = (sz_t * fzc_t + d.pd_fzl[z0 + st]) - (sz_b * fzc_b + d.pd_fzl[z0]);
    const real tend = R(0.0) + (((-dxi) * gx - dyi * gy) - d.rdnw[k] * gz);
    qdst[z0] = scalar_update(d, k, j * nx + i, q0[z0], tend, mup0, dt, 1);
}

// ---------------------------------------------------------------------------
// Scalar-transport tangent and adjoint (stage checkpoints)
// ---------------------------------------------------------------------------

// Post-acoustic mu' of the stage into a checkpoint slot.
__global__ void k_mup_new(Dom d, real* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < d.ny * d.nx) out[c] = mup_new_at(d, c);
}

// One stage of the unlimited transport with explicit checkpointed
// fluxes and masses: the arithmetic of k_scalar_update.
__global__ void k_tracer_stage(Dom d, const real* __restrict__ qcur,
                               const real* __restrict__ q0,
                               real* __restrict__ qdst,
                               const real* __restrict__ ru,
                               const real* __restrict__ rv,
                               const real* __restrict__ ww,
                               const real* __restrict__ mup_t,
                               const real* __restrict__ mup_new, real dt,
                               int clamp, real rdx, real rdy)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    Arr3 Q{qcur, ny, nx}, RU{ru, ny, nx + 1}, RV{rv, ny + 1, nx},
        RW{ww, ny, nx};
    const real tend = R(0.0) + adv_scalar(Q, RU, RV, RW, k, j, i, nz, ny, nx,
                                         d.rdnw, d.fnm, d.fnp, rdx, rdy);
    const int x = iM(k, j, i, ny, nx), c = j * nx + i;
    const real mu0 = d.mub2d[c] + mup_t[c];
    const real mu = d.mub2d[c] + mup_new[c];
    const real chm0 = d.c1h[k] * mu0 + d.c2h[k];
    const real chm = d.c1h[k] * mu + d.c2h[k];
    real v = chm0 * q0[x];
    const real t = dt * tend;
    v = v + t;
    v = v / chm;
    if (clamp && !(v > R(0.0)) && v == v) v = R(0.0);
    qdst[x] = v;
}

// One scalar stage's transpose with the dynamics' share: for lam (adjoint
// of q_new) the per-point pieces -- the adjoint of the flux divergence's
// output (g_tend), the time-t scalar's (added to gq0) and the two column
// masses' per-level terms (P0: mu0, P1: mu_new) -- from the forward's own
// tendency (the clamp's mask and d q_new / d C(mu_new) need q_new).
__global__ void k_ss_adj_pts(Dom d, const real* __restrict__ q,
                             const real* __restrict__ q0,
                             const real* __restrict__ ru,
                             const real* __restrict__ rv,
                             const real* __restrict__ ww,
                             const real* __restrict__ mup_t,
                             const real* __restrict__ mup_new,
                             const real* __restrict__ lam,
                             real* __restrict__ gq0, real* __restrict__ wk,
                             real dt, int clamp, real rdx, real rdy)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    Arr3 Q{q, ny, nx}, RU{ru, ny, nx + 1}, RV{rv, ny + 1, nx}, RW{ww, ny, nx};
    const real tend = R(0.0) + adv_scalar(Q, RU, RV, RW, k, j, i, nz, ny, nx,
                                         d.rdnw, d.fnm, d.fnp, rdx, rdy);
    const int x = iM(k, j, i, ny, nx), c = j * nx + i;
    const size_t n = (size_t)nz * ny * nx;
    const real mu0 = d.mub2d[c] + mup_t[c];
    const real mu = d.mub2d[c] + mup_new[c];
    const real chm0 = d.c1h[k] * mu0 + d.c2h[k];
    const real chm = d.c1h[k] * mu + d.c2h[k];
    real v = chm0 * q0[x];
    const real t = dt * tend;
    v = v + t;
    v = v / chm;
    const real lv = (clamp && !(v > R(0.0))) ? R(0.0) : lam[x];
    wk[x] = dt * lv / chm;
    gq0[x] += chm0 * lv / chm;
    wk[n + x] = d.c1h[k] * q0[x] * lv / chm;
    wk[2 * n + x] = -(d.c1h[k] * v * lv / chm);
}

__global__ __launch_bounds__(64 * KG) void k_ss_adj_cols(Dom d,
                                                          const real* wk,
                                                          real* gmu0,
                                                          real* gmu)
{
    const int nz = d.nz, st = d.ny * d.nx;
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const bool on = c < st;
    const size_t n = (size_t)nz * st;
    real v[2] = {R(0.0), R(0.0)};
    if (on)
        for (int k = ty; k < nz; k += KG) {
            v[0] += wk[n + (size_t)k * st + c];
            v[1] += wk[2 * n + (size_t)k * st + c];
        }
    group_sums(v);
    if (!on || ty != 0) return;
    gmu0[c] += v[0];
    gmu[c] += v[1];
}

// Adjoint stage, part 1: m = g / C(mu_new) (the adjoint of the division)
// and gq (+)= C(mu_t) m (the adjoint of the time-t term).
__global__ void k_tracer_adj_scale(Dom d, const real* __restrict__ g,
                                   real* __restrict__ m, real* __restrict__ gq,
                                   const real* __restrict__ mup_t,
                                   const real* __restrict__ mup_new, int first)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int x = iM(k, j, i, ny, nx), c = j * nx + i;
    const real chm0 = d.c1h[k] * (d.mub2d[c] + mup_t[c]) + d.c2h[k];
    const real chm = d.c1h[k] * (d.mub2d[c] + mup_new[c]) + d.c2h[k];
    const real mm = g[x] / chm;
    m[x] = mm;
    gq[x] = (first ? R(0.0) : gq[x]) + chm0 * mm;
}

// flux5 / flux3 coefficient of the o-th stencil cell (flux5: qm3..qp2,
// flux3: qm2..qp1), for the transposed gathers.
__device__ __forceinline__ real w5(int o, real v)
{
    const real a = fabs(v);
    switch (o) {
    case 0: return (v + a) / R(60.0);
    case 1: return (-R(8.0) * v - R(5.0) * a) / R(60.0);
    case 2: return (R(37.0) * v + R(10.0) * a) / R(60.0);
    case 3: return (R(37.0) * v - R(10.0) * a) / R(60.0);
    case 4: return (-R(8.0) * v + R(5.0) * a) / R(60.0);
    default: return (v - a) / R(60.0);
    }
}

__device__ __forceinline__ real w3(int o, real v)
{
    const real a = fabs(v);
    switch (o) {
    case 0: return (-v + a) / R(12.0);
    case 1: return (R(7.0) * v - R(3.0) * a) / R(12.0);
    case 2: return (R(7.0) * v + R(3.0) * a) / R(12.0);
    default: return (-v - a) / R(12.0);
    }
}

// Adjoint stage, part 2: gout = dt * D^T m, D the flux-form divergence
// of flux_div_scalar.  Face f (between cells f-1 and f) carries the
// adjoint Lambda_f = rdx (m[f] - m[f-1]); cell c gathers it with the
// weight it had in face f's stencil.
__global__ void k_tracer_adj_div(Dom d, const real* __restrict__ m,
                                 real* __restrict__ gout,
                                 const real* __restrict__ ru,
                                 const real* __restrict__ rv,
                                 const real* __restrict__ ww, real dt,
                                 real rdx, real rdy)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    auto M = [&](int kk, int jj, int ii) { return m[iM(kk, jj, ii, ny, nx)]; };
    real acc = R(0.0);
    for (int o = 0; o < 6; ++o) {
        // x face f reads cell f + o - 3 == i
        const int f = pmod(i + 3 - o, nx);
        const real lx = rdx * (M(k, j, f) - M(k, j, pmod(f - 1, nx)));
        acc += w5(o, ru[iU(k, j, f, ny, nx)]) * lx;
        const int g = pmod(j + 3 - o, ny);
        const real ly = rdy * (M(k, g, i) - M(k, pmod(g - 1, ny), i));
        acc += w5(o, rv[iV(k, g, i, ny, nx)]) * ly;
    }
    // eta faces kf = 1..nz-1 touching cell k
    for (int kf = k - 1; kf <= k + 2; ++kf) {
        if (kf < 1 || kf > nz - 1) continue;
        const real lz = d.rdnw[kf] * M(kf, j, i) - d.rdnw[kf - 1] * M(kf - 1, j, i);
        const real vel = ww[iM(kf, j, i, ny, nx)];
        if (kf == 1 || kf == nz - 1) {
            if (k == kf) acc += vel * d.fnm[kf] * lz;
            else if (k == kf - 1) acc += vel * d.fnp[kf] * lz;
        } else {
            acc += w3(k - kf + 2, vel) * lz;   // face kf reads cell kf+o-2
        }
    }
    gout[iM(k, j, i, ny, nx)] = dt * acc;
}

// ---------------------------------------------------------------------------
// Transpose of the final stage's PD-limited update (reference
// pd_advection.cu pd_fluxes -> pd_scale -> pd_renorm_apply, moist.py:
// 385-398).  The forward is re-run by k_pd_fluxes / k_pd_scale / k_pd_apply
// on a domain copy whose ru_m/rv_m/ww_m are the given fluxes, mup the
// post-stage mu' (mu_pp a zero field) and pd_* work slices.  Then, all in
// gather form:
//   k_pda_update  the clamped coupled update -> tg (adjoint of the
//                 limited divergence) and the time-t / mass shares;
//   k_pda_scale   each cell's factor adjoint, gathered from the correction
//                 fluxes it donates to, through the factor's clamps to the
//                 upwind budget (ph_low) and the outgoing correction (fo);
//   k_pda_faces   per face: the correction flux's and the upwind flux's
//                 adjoints, the velocity gradient, the upwind flux's q0
//                 shares and its face-mass adjoint;
//   k_pda_gather  per cell: the high-order flux transpose to q, the upwind
//                 shares to q0, the face masses to the column mass;
//   k_ss_adj_cols the two column-mass sums.
// Non-smooth points are differentiated as the branch taken: the clamp at
// zero, each correction's donor (the flux sign), the factor's [0, 1]
// clamps, the upwind flux's CFL caps; |cr| as sign(cr) cr.
// ---------------------------------------------------------------------------

struct PdWork {
    real *fxl, *fxc, *fyl, *fyc, *fzl, *fzc, *sc, *qn, *zero;
    real *tg;                    // tg, then P0 (mu_t), P1 (mu_new), gdl, gfo
    real *hx, *bx, *ax, *mx;     // x faces: correction adj, q0 shares, mass
    real *hy, *by, *ay, *my;
    real *hz, *bz, *az, *mz;
};

__device__ __forceinline__ real sgnr(real v)
{
    return v > R(0.0) ? R(1.0) : (v < R(0.0) ? -R(1.0) : R(0.0));
}

__global__ void k_pda_update(Dom d, PdWork w, const real* __restrict__ q0,
                             const real* __restrict__ mup_t,
                             const real* __restrict__ mup_new,
                             const real* __restrict__ lam,
                             real* __restrict__ gq0, real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int x = iM(k, j, i, ny, nx), c = j * nx + i;
    const size_t n = (size_t)nz * ny * nx;
    const real chm0 = d.c1h[k] * (d.mub2d[c] + mup_t[c]) + d.c2h[k];
    const real chm = d.c1h[k] * (d.mub2d[c] + mup_new[c]) + d.c2h[k];
    const real v = w.qn[x];
    const real lv = (v > R(0.0)) ? lam[x] : R(0.0);
    w.tg[x] = dt * lv / chm;
    gq0[x] += chm0 * lv / chm;
    w.tg[n + x] = d.c1h[k] * q0[x] * lv / chm;
    w.tg[2 * n + x] = -(d.c1h[k] * v * lv / chm);
}

__global__ void k_pda_scale(Dom d, PdWork w, const real* __restrict__ q0,
                            const real* __restrict__ mup_t,
                            real* __restrict__ gq0, real dxi, real dyi,
                            real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t n = (size_t)nz * st;
    const int z0 = iM(k, j, i, ny, nx), c = j * nx + i;
    const int u0 = iU(k, j, i, ny, nx), v0 = iV(k, j, i, ny, nx);
    const real* tg = w.tg;
    // face jumps of tg: x face f sees cells f-1 (if f >= 1) and f (if f < nx)
    auto jx = [&](int f) {
        const real r = f < nx ? tg[iM(k, j, f, ny, nx)] : R(0.0);
        const real l = f >= 1 ? tg[iM(k, j, f - 1, ny, nx)] : R(0.0);
        return dxi * (r - l);
    };
    auto jy = [&](int g) {
        const real r = g < ny ? tg[iM(k, g, i, ny, nx)] : R(0.0);
        const real l = g >= 1 ? tg[iM(k, g - 1, i, ny, nx)] : R(0.0);
        return dyi * (r - l);
    };
    auto jz = [&](int kf) {
        const real u = kf < nz ? d.rdnw[kf] * tg[iM(kf, j, i, ny, nx)] : R(0.0);
        const real b = kf >= 1 ? d.rdnw[kf - 1] * tg[iM(kf - 1, j, i, ny, nx)]
                               : R(0.0);
        return u - b;
    };
    // this cell's factor adjoint: the faces it donates to
    real gs = R(0.0);
    {
        const real fl = w.fxc[u0], fr = w.fxc[u0 + 1];
        if (fl < R(0.0)) gs += fl * jx(i);
        if (fr > R(0.0)) gs += fr * jx(i + 1);
        if (i == nx - 1) {
            const real f0 = w.fxc[iU(k, j, 0, ny, nx)];
            if (f0 > R(0.0)) gs += f0 * jx(0);
        }
        if (i == 0) {
            const real fn = w.fxc[iU(k, j, nx, ny, nx)];
            if (fn < R(0.0)) gs += fn * jx(nx);
        }
    }
    {
        const real fl = w.fyc[v0], fr = w.fyc[v0 + nx];
        if (fl < R(0.0)) gs += fl * jy(j);
        if (fr > R(0.0)) gs += fr * jy(j + 1);
        if (j == ny - 1) {
            const real f0 = w.fyc[iV(k, 0, i, ny, nx)];
            if (f0 > R(0.0)) gs += f0 * jy(0);
        }
        if (j == 0) {
            const real fn = w.fyc[iV(k, ny, i, ny, nx)];
            if (fn < R(0.0)) gs += fn * jy(ny);
        }
    }
    {
        const real fb = w.fzc[z0], ft = w.fzc[z0 + st];
        if (fb > R(0.0)) gs += fb * jz(k);
        if (ft < R(0.0)) gs += ft * jz(k + 1);
    }
    // the factor: min(1, max(0, ph_low / (fo + 1e-20))) where fo > ph_low
    const real chm0 = d.c1h[k] * (d.mub2d[c] + mup_t[c]) + d.c2h[k];
    const real dl = (dxi * (w.fxl[u0 + 1] - w.fxl[u0])
                      + dyi * (w.fyl[v0 + nx] - w.fyl[v0]))
                     + d.rdnw[k] * (w.fzl[z0 + st] - w.fzl[z0]);
    const real ph_low = chm0 * q0[z0] - dt * dl;
    const real fo = dt * ((dxi * (fmax(R(0.0), w.fxc[u0 + 1])
                                   - fmin(R(0.0), w.fxc[u0]))
                            + dyi * (fmax(R(0.0), w.fyc[v0 + nx])
                                     - fmin(R(0.0), w.fyc[v0])))
                           + d.rdnw[k] * (fmin(R(0.0), w.fzc[z0 + st])
                                          - fmax(R(0.0), w.fzc[z0])));
    real gpl = R(0.0), gfo = R(0.0);
    if (fo > ph_low) {
        const real den = fo + R(1e-20);
        const real ratio = ph_low / den;
        if (ratio > R(0.0) && ratio < R(1.0)) {
            gpl = gs / den;
            gfo = -(gs * ratio / den);
        }
    }
    gq0[z0] += chm0 * gpl;
    w.tg[n + z0] += d.c1h[k] * q0[z0] * gpl;
    w.tg[3 * n + z0] = -(dt * gpl);
    w.tg[4 * n + z0] = gfo;
}

// WRF flux_upwind's transpose: (dqm1, dq0v) coefficients and d/dcr.
__device__ __forceinline__ void upwind_adj(real qm1, real q0v, real cr,
                                           real g, real& dqm1, real& dq0v,
                                           real& dcr)
{
    const real a = cr + fabs(cr), b = cr - fabs(cr), sg = sgnr(cr);
    dqm1 = R(0.5) * fmin(R(1.0), a) * g;
    dq0v = R(0.5) * fmax(-R(1.0), b) * g;
    const real da = a < R(1.0) ? R(1.0) + sg : R(0.0);
    const real db = b > -R(1.0) ? R(1.0) - sg : R(0.0);
    dcr = R(0.5) * g * (qm1 * da + q0v * db);
}

__global__ void k_pda_faces(Dom d, PdWork w, const real* __restrict__ q,
                            const real* __restrict__ q0,
                            const real* __restrict__ ru,
                            const real* __restrict__ rv,
                            const real* __restrict__ ww,
                            const real* __restrict__ mup_new,
                            real* __restrict__ gru, real* __restrict__ grv,
                            real* __restrict__ gww, real dx, real dy,
                            real dxi, real dyi, real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const size_t n = (size_t)nz * ny * nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    auto mu = [&](int jj, int ii) { return d.mub2d[jj * nx + ii] + mup_new[jj * nx + ii]; };
    auto cellv = [&](const real* a, int kk, int jj, int ii) {
        return a[iM(kk, jj, ii, ny, nx)];
    };
    const real* tg = w.tg;
    const real* gdl = w.tg + 3 * n;
    const real* gfo = w.tg + 4 * n;
    auto d5 = [&](const real* x, real vel) {   // d flux5 / d vel
        return ((R(37.0) * (x[3] + x[2]) - R(8.0) * (x[4] + x[1]) + (x[5] + x[0]))
                - sgnr(vel) * (R(10.0) * (x[3] - x[2]) - R(5.0) * (x[4] - x[1])
                               + (x[5] - x[0]))) / R(60.0);
    };
    if (j < ny && k < nz) {                    // x face f = i
        const int f = i, a = (f >= nx) ? f - nx : f, b = (f == 0) ? nx - 1 : f - 1;
        const int xf = iU(k, j, f, ny, nx);
        const bool hr = f < nx, hl = f >= 1;
        const real fc = w.fxc[xf];
        const real jmp = dxi * ((hr ? cellv(tg, k, j, f) : R(0.0))
                                - (hl ? cellv(tg, k, j, f - 1) : R(0.0)));
        const real sf = fc > R(0.0) ? cellv(w.sc, k, j, b)
                        : (fc < R(0.0) ? cellv(w.sc, k, j, a) : R(1.0));
        real gc = sf * jmp;
        if (fc > R(0.0) && hl) gc += dt * dxi * cellv(gfo, k, j, f - 1);
        if (fc < R(0.0) && hr) gc -= dt * dxi * cellv(gfo, k, j, f);
        const real glw = jmp + dxi * ((hl ? cellv(gdl, k, j, f - 1) : R(0.0))
                                      - (hr ? cellv(gdl, k, j, f) : R(0.0)));
        const real G = glw - gc;
        const real vel = ru[xf];
        const real muf = (d.c1h[k] * R(0.5)) * (mu(j, a) + mu(j, b)) + d.c2h[k];
        const real cr = ((vel * dt) / dx) / muf;
        const real qb = cellv(q0, k, j, b), qa = cellv(q0, k, j, a);
        const real fu = flux_upwind(qb, qa, cr);
        real dqb, dqa, dcr;
        upwind_adj(qb, qa, cr, muf * (dx / dt) * G, dqb, dqa, dcr);
        real x[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) x[o] = cellv(q, k, j, pmod(f + o - 3, nx));
        gru[xf] += d5(x, vel) * gc + dcr * (dt / dx) / muf;
        w.hx[xf] = gc;
        w.bx[xf] = dqb;
        w.ax[xf] = dqa;
        w.mx[xf] = (dx / dt) * fu * G - dcr * cr / muf;
    }
    if (i < nx && k < nz) {                    // y face g = j
        const int g = j, a = (g >= ny) ? g - ny : g, b = (g == 0) ? ny - 1 : g - 1;
        const int yf = iV(k, g, i, ny, nx);
        const bool hr = g < ny, hl = g >= 1;
        const real fc = w.fyc[yf];
        const real jmp = dyi * ((hr ? cellv(tg, k, g, i) : R(0.0))
                                - (hl ? cellv(tg, k, g - 1, i) : R(0.0)));
        const real sf = fc > R(0.0) ? cellv(w.sc, k, b, i)
                        : (fc < R(0.0) ? cellv(w.sc, k, a, i) : R(1.0));
        real gc = sf * jmp;
        if (fc > R(0.0) && hl) gc += dt * dyi * cellv(gfo, k, g - 1, i);
        if (fc < R(0.0) && hr) gc -= dt * dyi * cellv(gfo, k, g, i);
        const real glw = jmp + dyi * ((hl ? cellv(gdl, k, g - 1, i) : R(0.0))
                                      - (hr ? cellv(gdl, k, g, i) : R(0.0)));
        const real G = glw - gc;
        const real vel = rv[yf];
        const real muf = (d.c1h[k] * R(0.5)) * (mu(a, i) + mu(b, i)) + d.c2h[k];
        const real cr = ((vel * dt) / dy) / muf;
        const real qb = cellv(q0, k, b, i), qa = cellv(q0, k, a, i);
        const real fu = flux_upwind(qb, qa, cr);
        real dqb, dqa, dcr;
        upwind_adj(qb, qa, cr, muf * (dy / dt) * G, dqb, dqa, dcr);
        real x[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) x[o] = cellv(q, k, pmod(g + o - 3, ny), i);
        grv[yf] += d5(x, vel) * gc + dcr * (dt / dy) / muf;
        w.hy[yf] = gc;
        w.by[yf] = dqb;
        w.ay[yf] = dqa;
        w.my[yf] = (dy / dt) * fu * G - dcr * cr / muf;
    }
    if (i < nx && j < ny) {                    // eta face kf = k
        const int kf = k, zf = iM(kf, j, i, ny, nx);
        real gc = R(0.0), dqb = R(0.0), dqa = R(0.0), gm = R(0.0);
        if (kf > 0 && kf < nz) {
            const real fc = w.fzc[zf];
            const real jmp = d.rdnw[kf] * cellv(tg, kf, j, i)
                             - d.rdnw[kf - 1] * cellv(tg, kf - 1, j, i);
            const real sf = fc > R(0.0) ? cellv(w.sc, kf, j, i)
                            : (fc < R(0.0) ? cellv(w.sc, kf - 1, j, i) : R(1.0));
            gc = sf * jmp;
            if (fc < R(0.0)) gc += dt * d.rdnw[kf - 1] * cellv(gfo, kf - 1, j, i);
            if (fc > R(0.0)) gc -= dt * d.rdnw[kf] * cellv(gfo, kf, j, i);
            const real glw = jmp + (d.rdnw[kf - 1] * cellv(gdl, kf - 1, j, i)
                                    - d.rdnw[kf] * cellv(gdl, kf, j, i));
            const real G = glw - gc;
            const real vel = ww[zf];
            const real dz = R(2.0) / (d.rdnw[kf] + d.rdnw[kf - 1]);
            const real muf = d.c1h[kf] * mu(j, i) + d.c2h[kf];
            const real cr = (vel * dt) / (dz * muf);
            const real qb = cellv(q0, kf - 1, j, i), qa = cellv(q0, kf, j, i);
            const real fu = flux_upwind(qb, qa, cr);
            real dcr;
            upwind_adj(qb, qa, cr, muf * (dz / dt) * G, dqb, dqa, dcr);
            real dv;
            if (kf == 1 || kf == nz - 1) {
                dv = d.fnm[kf] * cellv(q, kf, j, i) + d.fnp[kf] * cellv(q, kf - 1, j, i);
            } else {
                const real qm2 = cellv(q, kf - 2, j, i), qm1 = cellv(q, kf - 1, j, i);
                const real qc0 = cellv(q, kf, j, i), qp1 = cellv(q, kf + 1, j, i);
                dv = ((R(7.0) * (qc0 + qm1) - (qp1 + qm2))
                      + sgnr(vel) * (R(3.0) * (qc0 - qm1) - (qp1 - qm2))) / R(12.0);
            }
            gww[zf] += dv * gc + dcr * dt / (dz * muf);
            gm = d.c1h[kf] * ((dz / dt) * fu * G - dcr * cr / muf);
        }
        w.hz[zf] = gc;
        w.bz[zf] = dqb;
        w.az[zf] = dqa;
        w.mz[zf] = gm;
    }
}

__global__ void k_pda_gather(Dom d, PdWork w, const real* __restrict__ ru,
                             const real* __restrict__ rv,
                             const real* __restrict__ ww,
                             real* __restrict__ gq, real* __restrict__ gq0)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const size_t n = (size_t)nz * st;
    const int z0 = iM(k, j, i, ny, nx);
    real acc = R(0.0);
    for (int o = 0; o < 6; ++o) {
        // x face f reads cell f + o - 3 == i; face nx copies face 0's stencil
        const int f = pmod(i + 3 - o, nx);
        acc += w5(o, ru[iU(k, j, f, ny, nx)]) * w.hx[iU(k, j, f, ny, nx)];
        if (f == 0)
            acc += w5(o, ru[iU(k, j, nx, ny, nx)]) * w.hx[iU(k, j, nx, ny, nx)];
        const int g = pmod(j + 3 - o, ny);
        acc += w5(o, rv[iV(k, g, i, ny, nx)]) * w.hy[iV(k, g, i, ny, nx)];
        if (g == 0)
            acc += w5(o, rv[iV(k, ny, i, ny, nx)]) * w.hy[iV(k, ny, i, ny, nx)];
    }
    for (int kf = k - 1; kf <= k + 2; ++kf) {
        if (kf < 1 || kf > nz - 1) continue;
        const int zf = iM(kf, j, i, ny, nx);
        const real vel = ww[zf], h = w.hz[zf];
        if (kf == 1 || kf == nz - 1) {
            if (k == kf) acc += vel * d.fnm[kf] * h;
            else if (k == kf - 1) acc += vel * d.fnp[kf] * h;
        } else {
            acc += w3(k - kf + 2, vel) * h;
        }
    }
    gq[z0] = acc;
    // upwind fluxes of q0: cell i is the east cell of face i and the west
    // cell of face i+1 (periodic copies: face nx east = 0, face 0 west = nx-1)
    const int u0 = iU(k, j, i, ny, nx), v0 = iV(k, j, i, ny, nx);
    real g0 = w.ax[u0] + w.bx[u0 + 1] + w.ay[v0] + w.by[v0 + nx];
    real gm = w.mx[u0] + w.mx[u0 + 1] + w.my[v0] + w.my[v0 + nx];
    if (i == 0) {
        g0 += w.ax[iU(k, j, nx, ny, nx)];
        gm += w.mx[iU(k, j, nx, ny, nx)];
    }
    if (i == nx - 1) {
        g0 += w.bx[iU(k, j, 0, ny, nx)];
        gm += w.mx[iU(k, j, 0, ny, nx)];
    }
    if (j == 0) {
        g0 += w.ay[iV(k, ny, i, ny, nx)];
        gm += w.my[iV(k, ny, i, ny, nx)];
    }
    if (j == ny - 1) {
        g0 += w.by[iV(k, 0, i, ny, nx)];
        gm += w.my[iV(k, 0, i, ny, nx)];
    }
    real gmz = R(0.0);
    if (k >= 1) { g0 += w.az[z0]; gmz = w.mz[z0]; }
    if (k + 1 <= nz - 1) g0 += w.bz[z0 + st];
    gq0[z0] += g0;
    w.tg[2 * n + z0] += (d.c1h[k] * R(0.5)) * gm + gmz;
}

__global__ void k_add3(real* __restrict__ out, const real* __restrict__ a,
                       const real* __restrict__ b, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = a[t] + b[t];
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned absbits(real x)
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
            real dz = ((basev(d.phb, k + 1, c, st, d.base3d) + d.php[t + st])
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
    KID_OMEGA = 0, KID_TEND_FACES, KID_TEND_COLS, KID_UV, KID_COLUMN,
    KID_FINISH, KID_SET_W_SURFACE, KID_SCALARS, KID_DIFF6, KID_SMAG,
    KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {
    "k_omega", "k_tend_faces", "k_tend_cols", "k_uv", "k_column",
    "k_finish", "k_set_w_surface", "moist_scalars", "k_diff6",
    "smag_held"};

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
    // column kernels address each array with 32-bit byte offsets
    if (big >= (1LL << 30))
        return fail(ARW_E_ARG, "domain exceeds 2^30 points per field");
    const void* req[] = {d->u, d->v, d->w, d->thp, d->php, d->mup, d->p,
                         d->al, d->alt, d->u0, d->v0, d->w0, d->thp0,
                         d->php0, d->mup0, d->ru_t, d->rv_t, d->rw_t,
                         d->rth_t, d->rph_t, d->rmu_t, d->u_pp, d->v_pp,
                         d->w_pp, d->th_pp, d->ph_pp, d->mu_pp, d->p_pp,
                         d->p_pp_old, d->ww_pp, d->al_pp, d->thb, d->pb,
                         d->alb, d->phb, d->mub2d, d->ht, d->c1h, d->c2h,
                         d->c1f, d->c2f, d->dnw, d->rdnw, d->dn, d->rdn,
                         d->fnm, d->fnp, d->ww, d->gam, d->dmdt_rf,
                         d->mup_next};
    for (const void* p : req)
        if (!p) return fail(ARW_E_ARG, "null array in arw_domain");
    if (d->hypsometric_opt < 0 || d->hypsometric_opt > 2)
        return fail(ARW_E_ARG, "hypsometric_opt must be 1 or 2");
    if (d->hypsometric_opt == 2
        && (!d->c3h || !d->c4h || !d->c3f || !d->c4f))
        return fail(ARW_E_ARG, "hypsometric_opt = 2 needs c3h, c4h, c3f, "
                               "c4f");
    if (d->n_extra < 0 || d->n_extra > ARW_MAX_EXTRA || d->n_load < 0
        || d->n_load > 4 || d->n_load > d->n_extra)
        return fail(ARW_E_ARG, "need 0 <= n_load <= min(4, n_extra), "
                               "n_extra <= ARW_MAX_EXTRA");
    if ((d->n_extra || d->h_diabatic) && !d->qv)
        return fail(ARW_E_ARG, "extra species and h_diabatic need a moist "
                               "domain");
    for (int e = 0; e < d->n_extra; ++e)
        if (!d->qx[e] || !d->qx0[e] || !d->qxs[e])
            return fail(ARW_E_ARG, "null extra species array");
    if (d->qv) {
        const void* moist[] = {d->qc, d->qr, d->qv0, d->qc0, d->qr0, d->qsv,
                               d->qsc, d->qsr, d->ru_m, d->rv_m, d->ww_m,
                               d->pd_fxl, d->pd_fxc, d->pd_fyl, d->pd_fyc,
                               d->pd_fzl, d->pd_fzc, d->pd_sc};
        for (const void* p : moist)
            if (!p) return fail(ARW_E_ARG, "moist domain: null scalar or scratch array");
    }
    return ARW_OK;
}

// The precision form's extra arrays (the float32 build only).
int32_t check_prec(const arw_params* p, const Dom* d)
{
    if (ARW_F64 || !p || !p->precision_form) return ARW_OK;
    if (d->h_diabatic)
        return fail(ARW_E_UNSUPPORTED, "the precision form carries no "
                                       "h_diabatic removal");
    if (d->hypsometric_opt == 2)
        return fail(ARW_E_UNSUPPORTED, "the precision form's EOS is the "
                                       "hypsometric_opt = 1 one");
    if (!d->tri64)
        return fail(ARW_E_ARG, "precision form needs the FP64 tridiagonal scratch tri64");
    if (!d->pp)
        return fail(ARW_E_ARG, "precision form needs the perturbation pressure array pp");
    return ARW_OK;
}

int32_t check_params(const arw_params* p, const Dom* d)
{
    if (!p) return fail(ARW_E_ARG, "null params");
    if (int32_t rc = check_prec(p, d)) return rc;
    if (!(p->dx > 0.0) || !(p->dy > 0.0) || !(p->dt > 0.0))
        return fail(ARW_E_ARG, "dx, dy, dt must be positive");
    if (p->time_step_sound < 1 || p->time_step_sound % 2)
        return fail(ARW_E_ARG, "time_step_sound must be even and positive");
    if ((p->khdif > 0.0 || p->kvdif > 0.0)
        && (!d->rdzf_h || !d->rdzc_h || !d->rdzf_w || !d->rdzc_w))
        return fail(ARW_E_ARG, "diffusion spacings required when khdif/kvdif > 0");
    if (p->moist_cq) {
        if (!d->qv) return fail(ARW_E_ARG, "moist_cq needs a moist domain");
        if (p->moist_cq < 1 || p->moist_cq > 3 + d->n_load
            || (p->moist_cq > 1 && p->moist_cq < 3))
            return fail(ARW_E_ARG, "moist_cq: 1, 3 or 3 + n_load species");
        if (!d->cqu || !d->cqv || !d->cqw)
            return fail(ARW_E_ARG, "moist_cq needs the cqu, cqv, cqw arrays");
        if (!ARW_F64 && p->precision_form)
            return fail(ARW_E_UNSUPPORTED, "moist_cq in the plain form only");
    }
    if (p->km_opt == 4) {
        if (!ARW_F64 && p->precision_form)
            return fail(ARW_E_UNSUPPORTED, "km_opt = 4: plain form only");
        if (d->qv && !d->held_q)
            return fail(ARW_E_ARG, "km_opt = 4 on a moist state needs held_q");
        if (d->nz < 3) return fail(ARW_E_ARG, "km_opt = 4 needs nz >= 3");
        if (!d->sm_buf || !d->d6_u || !d->d6_v || !d->d6_w || !d->d6_t)
            return fail(ARW_E_ARG, "km_opt = 4 needs sm_buf and the d6_* "
                                   "held buffers");
        if (p->isfflx < 0 || p->isfflx > 2)
            return fail(ARW_E_ARG, "isfflx must be 0, 1 or 2");
        if (!(p->c_s >= 0.0)) return fail(ARW_E_ARG, "c_s must be >= 0");
    }
    return ARW_OK;
}

// The stage-piece entry points take the stage cq as arw_step computes it
// internally; they run with moist_cq = 0 only.
static int32_t refuse_cq(const arw_params* p)
{
    if (p && p->moist_cq)
        return fail(ARW_E_UNSUPPORTED, "moist_cq: arw_step forms the stage "
                                       "cq itself; the stage-piece entry "
                                       "points take moist_cq = 0");
    return ARW_OK;
}

inline int32_t launched(hipError_t e, const char* what, int kid,
                        hipStream_t st)
{
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
        return ARW_E_LAUNCH;
    }
    if (kid >= 0) prof_mark(kid, st);
    return ARW_OK;
}

inline int32_t launched(const char* what)
{
    return launched(hipGetLastError(), what, -1, nullptr);
}

inline dim3 col_grid(const Dom& d) { return dim3((d.ny * d.nx + 255) / 256); }
inline dim3 union_grid(const Dom& d, int nlev)
{
    return dim3((d.nx + 1 + 63) / 64, (d.ny + 1 + 3) / 4, nlev);
}

// Stage preparation: Omega (+ stage EOS), slow tendencies, small-step
// initialisation.
int32_t stage_prepare(const Dom& d, const Sc& s, int do_eos, hipStream_t st)
{
    int32_t rc = launched(run_omega(d, s, do_eos, st), "k_omega", KID_OMEGA, st);
    if (rc) return rc;
    rc = launched(run_tend_faces(d, s, st), "k_tend_faces", KID_TEND_FACES, st);
    if (rc) return rc;
    return launched(run_tend_cols(d, s, st), "k_tend_cols", KID_TEND_COLS, st);
}

int32_t substep(const Dom& d, const Sc& s, real* pout, int write_al,
                int write_ww, hipStream_t st)
{
    int32_t rc = launched(run_uv(d, s, st), "k_uv", KID_UV, st);
    if (rc) return rc;
    return launched(run_column(d, s, pout, write_al, write_ww, st), "k_column",
                    KID_COLUMN, st);
}

// Below this many columns the column marches leave most SIMDs with less
// than four waves, and the thread-per-point scalar kernels are faster
// (both are bit-identical).
constexpr long kScalarColumnsMin = 512L * 512;

// Every scalar's RK stage (reference moist.py:434-698): the unlimited
// flux form, or the positive-definite limiter on the final stage when
// moist_adv_opt = 1.  cur: stage-state scalars; q0: time t; dst: new.
// The limited update reads q0's neighbours while it runs, so it goes to
// spare (a dead buffer of the stage rotation) and is copied to dst.
int32_t scalar_stage(const Dom& dd, const arw_params* prm, real* const* cur,
                     real* const* q0_in, real* const* dst, const real* mup0,
                     double dt_stage, bool final, hipStream_t st, int nq,
                     real* const* spare, real* const* fixed)
{
    const real dt = (real)dt_stage;
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const bool pd = final && prm->moist_adv_opt == 1;
    int32_t rc;
    const real* q0[3 + ARW_MAX_EXTRA];
    const real* fx[3 + ARW_MAX_EXTRA];
    for (int q = 0; q < nq; ++q) {
        q0[q] = q0_in[q];
        fx[q] = fixed ? fixed[q] : nullptr;
        if (pd && fx[q]) {
            // the limiter advects q_t + dt fixed / C(mu_t): fold the held
            // tendency into the time-t scalar in its own slot (its last use)
            hipLaunchKernelGGL(k_pd_fold, dim3((dd.nx + 63) / 64,
                                               (dd.ny + 3) / 4, dd.nz),
                               dim3(64, 4), 0, st, dd, q0_in[q], fixed[q],
                               mup0, dt);
            if ((rc = launched(hipGetLastError(), "k_pd_fold", KID_SCALARS,
                               st)))
                return rc;
            q0[q] = fixed[q];
            fx[q] = nullptr;
        }
    }
    if ((long)dd.ny * dd.nx < kScalarColumnsMin) {
        // few columns: one thread per point (stored limiter fluxes)
        const dim3 cells((dd.nx + 63) / 64, (dd.ny + 3) / 4, dd.nz);
        for (int q = 0; q < nq; ++q) {
            if (pd) {
                hipLaunchKernelGGL(k_pd_fluxes, union_grid(dd, dd.nz + 1),
                                   dim3(64, 4), 0, st, dd, cur[q], q0[q],
                                   (real)prm->dx, (real)prm->dy, dt);
                hipLaunchKernelGGL(k_pd_scale, cells, dim3(64, 4), 0, st, dd,
                                   q0[q], mup0, rdx, rdy, dt);
                hipLaunchKernelGGL(k_pd_apply, cells, dim3(64, 4), 0, st, dd,
                                   q0[q], dst[q], mup0, rdx, rdy, dt);
            } else {
                hipLaunchKernelGGL(k_scalar_update, cells, dim3(64, 4), 0, st,
                                   dd, cur[q], q0[q], dst[q], mup0, dt,
                                   final ? 1 : 0, rdx, rdy, fx[q]);
            }
            if ((rc = launched(hipGetLastError(), "scalar stage", KID_SCALARS,
                               st)))
                return rc;
        }
        return ARW_OK;
    }
    if (pd) {
        const size_t bytes = (size_t)dd.nz * dd.ny * dd.nx * sizeof(real);
        for (int q = 0; q < nq; ++q) {
            if ((rc = launched(run_pd_cols(dd, cur[q], q0[q], spare[q], mup0,
                                           (real)prm->dx, (real)prm->dy, rdx,
                                           rdy, dt, st),
                               "k_pd_cols", KID_SCALARS, st)))
                return rc;
            if (hipMemcpyAsync(dst[q], spare[q], bytes,
                               hipMemcpyDeviceToDevice, st) != hipSuccess)
                return fail(ARW_E_LAUNCH, "limited scalar copy failed");
        }
        return ARW_OK;
    }
    for (int q = 0; q < nq; q += 4) {
        QGroup g;
        g.n = nq - q < 4 ? nq - q : 4;
        for (int e = 0; e < 4; ++e) {
            const int k = e < g.n ? q + e : q;
            g.cur[e] = cur[k];
            g.q0[e] = q0[k];
            g.dst[e] = dst[k];
            g.fixed[e] = fx[k];
        }
        if ((rc = launched(run_scalar_cols(dd, g, mup0, dt, rdx, rdy,
                                           final ? 1 : 0, st),
                           "k_scalar_cols", KID_SCALARS, st)))
            return rc;
    }
    return ARW_OK;
}

// ---------------------------------------------------------------------------
// Ensemble launch sequences
// ---------------------------------------------------------------------------

// Member views: per stage s the stage view (role 3s) and the substep views
// with p''/p''_old in place (3s+1) and swapped (3s+2), then the time-level
// view for the lower boundary (ENS_BASE).
enum { ENS_BASE = 9, ENS_ROLES = 10 };

struct EnsGroup {
    BatchSet sets[ENS_ROLES];
};

// The views arw_step's stage loop forms for one member (same rotation).
void member_views(const Dom& d, Dom* role, FinishDst* fin)
{
    real* A[6] = {d.u, d.v, d.w, d.thp, d.php, d.mup};
    real* B[6] = {d.u0, d.v0, d.w0, d.thp0, d.php0, d.mup0};
    real* mup_c = d.mup_next;
    for (int stage = 0; stage < 3; ++stage) {
        Dom dd = d;
        dd.u0 = A[0]; dd.v0 = A[1]; dd.w0 = A[2]; dd.thp0 = A[3];
        dd.php0 = A[4]; dd.mup0 = A[5];
        real* const* T = (stage == 0) ? A : B;
        dd.u = T[0]; dd.v = T[1]; dd.w = T[2]; dd.thp = T[3]; dd.php = T[4];
        dd.mup = (stage == 0) ? A[5] : (stage == 1 ? B[5] : mup_c);
        real* const* D = (stage == 2) ? A : B;
        FinishDst& o = fin[stage];
        o.u = D[0]; o.v = D[1]; o.w = D[2]; o.thp = D[3]; o.php = D[4];
        o.mup = (stage == 0) ? B[5] : (stage == 1 ? mup_c : A[5]);
        role[3 * stage] = dd;
        role[3 * stage + 1] = dd;
        Dom sw = dd;
        sw.p_pp = dd.p_pp_old;
        sw.p_pp_old = dd.p_pp;
        role[3 * stage + 2] = sw;
    }
    role[ENS_BASE] = d;
}

// One step of one group: the launch sequence of arw_step, each launch
// covering every member of the group.
int32_t ens_issue(const EnsGroup& G, const arw_params* prm, bool b3,
                  hipStream_t st)
{
    const int ns = prm->time_step_sound;
    const int nsub[3] = {1, ns / 2 > 1 ? ns / 2 : 1, ns};
    const double dtaus[3] = {prm->dt / 3.0, prm->dt / ns, prm->dt / ns};
    int32_t rc;
    for (int stage = 0; stage < 3; ++stage) {
        const BatchSet& DD = G.sets[3 * stage];
        Sc s = make_sc(prm, dtaus[stage], 1);
        if ((rc = launched(run_omega_b(DD, s, stage == 0, b3, st), "k_omega (ensemble)", -1, st)))
            return rc;
        if ((rc = launched(run_tend_faces_b(DD, s, b3, st), "k_tend_faces (ensemble)", -1, st)))
            return rc;
        if ((rc = launched(run_tend_cols_b(DD, s, b3, st), "k_tend_cols (ensemble)", -1, st)))
            return rc;
        for (int n = 0; n < nsub[stage]; ++n) {
            Sc sn = make_sc(prm, dtaus[stage], n == 0);
            sn.use_pold = (n > 0);
            sn.gam_ready = (n > 0);
            const bool last = (n + 1 == nsub[stage]);
            const BatchSet& DS = G.sets[3 * stage + 1 + (n & 1)];
            if ((rc = launched(run_uv_b(DS, sn, b3, st), "k_uv (ensemble)", -1, st)))
                return rc;
            if ((rc = launched(run_column_b(DS, sn, b3, b3 || last, st), "k_column (ensemble)", -1, st)))
                return rc;
        }
        if ((rc = launched(run_finish_b(DD, b3, s.prec != 0, st), "k_finish (ensemble)", -1, st)))
            return rc;
    }
    Sc s = make_sc(prm, prm->dt, 1);
    const Batch& bc = G.sets[ENS_BASE].k[BK_COLS];
    hipLaunchKernelGGL(k_set_w_surface<Many>, dim3(xcd_grid((unsigned)bc.total)),
                       dim3(256), 0, st, bc, s);
    return launched(hipGetLastError(), "k_set_w_surface (ensemble)", -1, st);
}

}  // namespace

int32_t lib_fail(int32_t code, const char* msg) { return fail(code, msg); }
int32_t lib_launched(hipError_t e, const char* what)
{
    return launched(e, what, -1, nullptr);
}
int32_t lib_check_domain(const Dom* d) { return check_domain(d); }
// The adjoint kernels differentiate the qv/qc/qr step without retained
// microphysics heating.
int32_t lib_check_adj_domain(const Dom* d)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (d->n_extra || d->h_diabatic)
        return fail(ARW_E_UNSUPPORTED, "the adjoint covers qv/qc/qr states "
                                       "without h_diabatic");
    if (d->hypsometric_opt == 2)
        return fail(ARW_E_UNSUPPORTED, "the adjoint covers hypsometric_opt "
                                       "= 1");
    return ARW_OK;
}
int32_t lib_check_prec(const arw_params* p, const Dom* d)
{
    return check_prec(p, d);
}

// Work layout of the PD stage entries (elements of real): the six face
// flux arrays, the factors, the recomputed q_new and a zero mu'' field
// (forward), then tg and four cell adjoints, then four adjoint arrays per
// face direction.  6 (X + Y + Z) + 7 n + ny nx <= 26 (nz+1)(ny+1)(nx+1).
PdWork pd_layout(const Dom& d, real* work)
{
    const size_t nz = d.nz, ny = d.ny, nx = d.nx;
    const size_t X = nz * ny * (nx + 1), Y = nz * (ny + 1) * nx;
    const size_t Z = (nz + 1) * ny * nx, n = nz * ny * nx;
    real* p = work;
    auto take = [&](size_t m) { real* r = p; p += m; return r; };
    PdWork w;
    w.fxl = take(X); w.fxc = take(X);
    w.fyl = take(Y); w.fyc = take(Y);
    w.fzl = take(Z); w.fzc = take(Z);
    w.sc = take(n); w.qn = take(n); w.zero = take(ny * nx);
    w.tg = take(5 * n);
    w.hx = take(X); w.bx = take(X); w.ax = take(X); w.mx = take(X);
    w.hy = take(Y); w.by = take(Y); w.ay = take(Y); w.my = take(Y);
    w.hz = take(Z); w.bz = take(Z); w.az = take(Z); w.mz = take(Z);
    return w;
}

__global__ void k_zero_real(real* p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = R(0.0);
}

// The fused step's PD kernels on explicit inputs: a domain copy reads
// the fluxes as its ru_m/rv_m/ww_m and the post-stage mu' as mup (with a
// zero mu''; x + 0 is exact), so q_new is bit-identical to the step's.
int32_t pd_forward(const Dom* d, const arw_params* prm, double dt_stage,
                   const real* q, const real* q0, const real* ru,
                   const real* rv, const real* ww, const real* mup_t,
                   const real* mup_new, real* q_new, const PdWork& w,
                   hipStream_t st)
{
    int32_t rc;
    Dom dd = *d;
    dd.mup = const_cast<real*>(mup_new);
    dd.mu_pp = w.zero;
    dd.ru_m = const_cast<real*>(ru);
    dd.rv_m = const_cast<real*>(rv);
    dd.ww_m = const_cast<real*>(ww);
    dd.pd_fxl = w.fxl; dd.pd_fxc = w.fxc;
    dd.pd_fyl = w.fyl; dd.pd_fyc = w.fyc;
    dd.pd_fzl = w.fzl; dd.pd_fzc = w.fzc;
    dd.pd_sc = w.sc;
    const int s2 = d->ny * d->nx;
    hipLaunchKernelGGL(k_zero_real, dim3((s2 + 255) / 256), dim3(256), 0, st,
                       w.zero, s2);
    if ((rc = launched("k_zero_real"))) return rc;
    const real dt = (real)dt_stage;
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const dim3 cells((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
    hipLaunchKernelGGL(k_pd_fluxes, union_grid(dd, dd.nz + 1), dim3(64, 4),
                       0, st, dd, q, q0, (real)prm->dx, (real)prm->dy, dt);
    if ((rc = launched("k_pd_fluxes"))) return rc;
    hipLaunchKernelGGL(k_pd_scale, cells, dim3(64, 4), 0, st, dd, q0, mup_t,
                       rdx, rdy, dt);
    if ((rc = launched("k_pd_scale"))) return rc;
    hipLaunchKernelGGL(k_pd_apply, cells, dim3(64, 4), 0, st, dd, q0, q_new,
                       mup_t, rdx, rdy, dt);
    return launched("k_pd_apply");
}

// WRF sixth_order_diffusion (reference kernels/diff6.cu, launched by
// dycore.py:1647-1740 from prepare_fixed_tendencies :1345-1420): the
// coupled tendency of a periodic field from Xue's (2000) eq. 3 face
// fluxes, the mono = 2 up-gradient knock-down, the face masses of the u
// (variant 1), v (variant 2) and mass/w positions and the
// diff_6th_slopeopt taper on a terrain base (msf = 1); out = 0 + (x part
// + y part), or added to a held Smagorinsky tendency already in out.
// The four rows of the step's own fields -- u, v, w (zero on the boundary
// levels) and theta' -- run in one launch over the union grid: each point
// gathers the column masses of its 3 x 3 neighbourhood and, tapered, the
// geopotential taps once for every row it owns.
__global__ void k_diff6_rows(Dom d, real coef, int mono, int slope,
                             real dzthr_x, real dzthr_y, int acc)
{
    const int ny = d.ny, nx = d.nx, nz = d.nz;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const bool du = j < ny && k < nz;               // u face (i <= nx)
    const bool dv = i < nx && k < nz;               // v face (j <= ny)
    const bool dw = i < nx && j < ny;               // w level (k <= nz)
    const bool dt = dw && k < nz;                   // mass point
    if (!(du || dv || dw)) return;
    const int ic = (i == nx) ? 0 : i, jc = (j == ny) ? 0 : j;
    const real ten = R(10.0), five = R(5.0), zero = R(0.0), one = R(1.0);
    const real q = R(0.25), h = R(0.5);
    real m[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int c = pmod(jc + a - 1, ny) * nx + pmod(ic + b - 1, nx);
            m[a][b] = d.mub2d[c] + d.mup[c];
        }
    // geopotential taps (jc + a - 2, ic + b - 2), a, b in 0..3
    real ph[4][4];
    const bool tap = slope && k < nz;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            ph[a][b] = tap ? d.phb[((size_t)k * ny + pmod(jc + a - 2, ny)) * nx
                                   + pmod(ic + b - 2, nx)]
                           : zero;
#define MU(dj, di) m[(dj) + 1][(di) + 1]
#define PHB(dj, di) ph[(dj) + 2][(di) + 2]
    auto row = [&](const real* f, real* out, size_t o, int nys, int nxs,
                   real cc1, real cc2, real mx0, real mx1, real my0, real my1,
                   real sx0, real sx1, real sy0, real sy1) {
        auto FX = [&](int mm) {
            return f[((size_t)k * nys + jc) * nxs + pmod(ic + mm, nx)];
        };
        auto FY = [&](int mm) {
            return f[((size_t)k * nys + pmod(jc + mm, ny)) * nxs + ic];
        };
        real px0 = ten * (FX(0) - FX(-1)) - five * (FX(1) - FX(-2)) + (FX(2) - FX(-3));
        real px1 = ten * (FX(1) - FX(0)) - five * (FX(2) - FX(-1)) + (FX(3) - FX(-2));
        if (mono == 2) {
            if (px0 * (FX(0) - FX(-1)) <= zero) px0 = zero;
            if (px1 * (FX(1) - FX(0)) <= zero) px1 = zero;
        }
        const real tx = coef * (sx1 * (cc1 * mx1 + cc2) * px1
                                - sx0 * (cc1 * mx0 + cc2) * px0);
        real py0 = ten * (FY(0) - FY(-1)) - five * (FY(1) - FY(-2)) + (FY(2) - FY(-3));
        real py1 = ten * (FY(1) - FY(0)) - five * (FY(2) - FY(-1)) + (FY(3) - FY(-2));
        if (mono == 2) {
            if (py0 * (FY(0) - FY(-1)) <= zero) py0 = zero;
            if (py1 * (FY(1) - FY(0)) <= zero) py1 = zero;
        }
        const real ty = coef * (sy1 * (cc1 * my1 + cc2) * py1
                                - sy0 * (cc1 * my0 + cc2) * py0);
        out[o] = acc ? out[o] + (zero + (tx + ty)) : zero + (tx + ty);
    };
    // the slope taper of a row: a0/a1 (b0/b1) the x (y) face steps, taken
    // with the variant's neighbouring steps (diff6.cu's d0/d1)
    const real a0 = fabs(PHB(0, 0) - PHB(0, -1));
    const real a1 = fabs(PHB(0, 1) - PHB(0, 0));
    const real b0 = fabs(PHB(0, 0) - PHB(-1, 0));
    const real b1 = fabs(PHB(1, 0) - PHB(0, 0));
    auto taper = [&](real dd, real thr) { return fmax(one - dd / thr, zero); };
    if (du) {
        real sx0 = one, sx1 = one, sy0 = one, sy1 = one;
        if (slope) {
            sx0 = taper(fmax(a0, fabs(PHB(0, -1) - PHB(0, -2))), dzthr_x);
            sx1 = taper(fmax(a1, a0), dzthr_x);
            sy0 = taper(fmax(b0, fabs(PHB(0, -1) - PHB(-1, -1))), dzthr_y);
            sy1 = taper(fmax(b1, fabs(PHB(1, -1) - PHB(0, -1))), dzthr_y);
        }
        row(d.u, d.d6_u, ((size_t)k * ny + j) * (nx + 1) + i, ny, nx + 1,
            d.c1h[k], d.c2h[k], MU(0, -1), MU(0, 0),
            q * (MU(-1, -1) + MU(-1, 0) + MU(0, -1) + MU(0, 0)),
            q * (MU(0, -1) + MU(0, 0) + MU(1, -1) + MU(1, 0)),
            sx0, sx1, sy0, sy1);
    }
    if (dv) {
        real sx0 = one, sx1 = one, sy0 = one, sy1 = one;
        if (slope) {
            sx0 = taper(fmax(a0, fabs(PHB(-1, 0) - PHB(-1, -1))), dzthr_x);
            sx1 = taper(fmax(a1, fabs(PHB(-1, 1) - PHB(-1, 0))), dzthr_x);
            sy0 = taper(fmax(b0, fabs(PHB(-1, 0) - PHB(-2, 0))), dzthr_y);
            sy1 = taper(fmax(b1, b0), dzthr_y);
        }
        row(d.v, d.d6_v, ((size_t)k * (ny + 1) + j) * nx + i, ny + 1, nx,
            d.c1h[k], d.c2h[k],
            q * (MU(-1, -1) + MU(-1, 0) + MU(0, -1) + MU(0, 0)),
            q * (MU(-1, 0) + MU(-1, 1) + MU(0, 0) + MU(0, 1)),
            MU(-1, 0), MU(0, 0), sx0, sx1, sy0, sy1);
    }
    if (dw || dt) {
        real sx0 = one, sx1 = one, sy0 = one, sy1 = one;
        if (slope) {
            sx0 = taper(a0, dzthr_x);
            sx1 = taper(a1, dzthr_x);
            sy0 = taper(b0, dzthr_y);
            sy1 = taper(b1, dzthr_y);
        }
        const real mx0 = h * (MU(0, -1) + MU(0, 0)), mx1 = h * (MU(0, 0) + MU(0, 1));
        const real my0 = h * (MU(-1, 0) + MU(0, 0)), my1 = h * (MU(0, 0) + MU(1, 0));
        const size_t o = ((size_t)k * ny + j) * nx + i;
        if (k == 0 || k == nz)
            d.d6_w[o] = acc ? d.d6_w[o] + R(0.0) : R(0.0);
        else
            row(d.w, d.d6_w, o, ny, nx, d.c1f[k], d.c2f[k], mx0, mx1, my0,
                my1, sx0, sx1, sy0, sy1);
        if (dt)
            row(d.thp, d.d6_t, o, ny, nx, d.c1h[k], d.c2h[k], mx0, mx1, my0,
                my1, sx0, sx1, sy0, sy1);
    }
#undef MU
#undef PHB
}

// The filter's species rows (mass points, plain periodic) in groups of
// up to four: the column masses, slope taper and C(mu) face factors of
// the point are formed once and shared; per species the arithmetic is
// k_diff6_rows' mass row, in its order.
struct D6Q {
    const real* f[4];
    real* out[4];
};

template <int NS>
__global__ void k_diff6q(Dom d, D6Q q, real coef, int mono, int slope,
                         real dzthr_x, real dzthr_y, int acc)
{
    const int ny = d.ny, nx = d.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= d.nz) return;
    const size_t o = ((size_t)k * ny + j) * nx + i;
    auto MU = [&](int jj, int ii) {
        const int c = pmod(jj, ny) * nx + pmod(ii, nx);
        return d.mub2d[c] + d.mup[c];
    };
    auto PHB = [&](int jj, int ii) {
        return d.phb[((size_t)k * ny + pmod(jj, ny)) * nx + pmod(ii, nx)];
    };
    const real ten = R(10.0), five = R(5.0), zero = R(0.0), one = R(1.0);
    const real mx0 = R(0.5) * (MU(j, i - 1) + MU(j, i));
    const real mx1 = R(0.5) * (MU(j, i) + MU(j, i + 1));
    const real my0 = R(0.5) * (MU(j - 1, i) + MU(j, i));
    const real my1 = R(0.5) * (MU(j, i) + MU(j + 1, i));
    real sx0 = one, sx1 = one, sy0 = one, sy1 = one;
    if (slope) {
        sx0 = fmax(one - fabs(PHB(j, i) - PHB(j, i - 1)) / dzthr_x, zero);
        sx1 = fmax(one - fabs(PHB(j, i + 1) - PHB(j, i)) / dzthr_x, zero);
        sy0 = fmax(one - fabs(PHB(j, i) - PHB(j - 1, i)) / dzthr_y, zero);
        sy1 = fmax(one - fabs(PHB(j + 1, i) - PHB(j, i)) / dzthr_y, zero);
    }
    const real c1 = d.c1h[k], c2 = d.c2h[k];
    const real cx1 = sx1 * (c1 * mx1 + c2), cx0 = sx0 * (c1 * mx0 + c2);
    const real cy1 = sy1 * (c1 * my1 + c2), cy0 = sy0 * (c1 * my0 + c2);
#pragma unroll
    for (int e = 0; e < NS; ++e) {
        const real* f = q.f[e];
        auto FX = [&](int m) { return f[((size_t)k * ny + j) * nx + pmod(i + m, nx)]; };
        auto FY = [&](int m) { return f[((size_t)k * ny + pmod(j + m, ny)) * nx + i]; };
        real px0 = ten * (FX(0) - FX(-1)) - five * (FX(1) - FX(-2)) + (FX(2) - FX(-3));
        real px1 = ten * (FX(1) - FX(0)) - five * (FX(2) - FX(-1)) + (FX(3) - FX(-2));
        real py0 = ten * (FY(0) - FY(-1)) - five * (FY(1) - FY(-2)) + (FY(2) - FY(-3));
        real py1 = ten * (FY(1) - FY(0)) - five * (FY(2) - FY(-1)) + (FY(3) - FY(-2));
        if (mono == 2) {
            if (px0 * (FX(0) - FX(-1)) <= zero) px0 = zero;
            if (px1 * (FX(1) - FX(0)) <= zero) px1 = zero;
            if (py0 * (FY(0) - FY(-1)) <= zero) py0 = zero;
            if (py1 * (FY(1) - FY(0)) <= zero) py1 = zero;
        }
        const real tx = coef * (cx1 * px1 - cx0 * px0);
        const real ty = coef * (cy1 * py1 - cy0 * py0);
        real* out = q.out[e];
        out[o] = acc ? out[o] + (zero + (tx + ty)) : zero + (tx + ty);
    }
}

// The four held tendencies from the time-t fields of d (u, v, w, thp and
// mub2d + mup), once per step.
int32_t held_diff6(const Dom& d, const arw_params* prm, hipStream_t st,
                   int acc, real* const* q_src, int nq)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const double clock_factor = prm->diff_6th_factor;
    const real coef = (real)(clock_factor * 0.015625 / (2.0 * prm->dt));
    const int slope = (prm->diff_6th_slopeopt >= 1 && d.base3d) ? 1 : 0;
    const real tx = (real)(prm->diff_6th_thresh * 9.81 * prm->dx);
    const real ty = (real)(prm->diff_6th_thresh * 9.81 * prm->dy);
    const int mono = prm->diff_6th_opt;
    hipLaunchKernelGGL(k_diff6_rows, dim3((nx + 1 + 63) / 64, (ny + 1 + 3) / 4,
                                          nz + 1),
                       dim3(64, 4), 0, st, d, coef, mono, slope, tx, ty, acc);
    if (int32_t rc = launched(hipGetLastError(), "k_diff6", KID_DIFF6, st))
        return rc;
    // the transported species the filter acts on (mass rows, reference
    // dycore.py:1387-1415), into their held slots
    const size_t n = (size_t)nz * ny * nx;
    int sel[3 + ARW_MAX_EXTRA], ns = 0;
    for (int e = 0; e < nq; ++e)
        if ((prm->diff6_species >> e) & 1u) sel[ns++] = e;
    const dim3 g((nx + 63) / 64, (ny + 3) / 4, nz), b(64, 4);
    for (int s0 = 0; s0 < ns; s0 += 4) {
        const int m = ns - s0 < 4 ? ns - s0 : 4;
        D6Q q;
        for (int e = 0; e < 4; ++e) {
            const int sp = sel[s0 + (e < m ? e : 0)];
            q.f[e] = q_src[sp];
            q.out[e] = d.held_q + sp * n;
        }
        switch (m) {
        case 4: hipLaunchKernelGGL(k_diff6q<4>, g, b, 0, st, d, q, coef, mono, slope, tx, ty, acc); break;
        case 3: hipLaunchKernelGGL(k_diff6q<3>, g, b, 0, st, d, q, coef, mono, slope, tx, ty, acc); break;
        case 2: hipLaunchKernelGGL(k_diff6q<2>, g, b, 0, st, d, q, coef, mono, slope, tx, ty, acc); break;
        default: hipLaunchKernelGGL(k_diff6q<1>, g, b, 0, st, d, q, coef, mono, slope, tx, ty, acc); break;
        }
        int32_t rc = launched(hipGetLastError(), "k_diff6", KID_DIFF6, st);
        if (rc) return rc;
    }
    return ARW_OK;
}

}  // namespace ARW_NS

using namespace ARW_NS;

struct ARW_API(ensemble) {
    int n = 0;
    arw_params prm{};
    bool b3 = false;
    void* dev = nullptr;                  // views and work tables
    std::vector<EnsGroup> groups;
    hipGraphExec_t exec = nullptr;
};

// ---------------------------------------------------------------------------
// The advection-only step (reference dycore.py:2318-2343, step(acoustic=
// False)): per stage of length dt/3, dt/2, dt the EOS, the coupled flux-
// form advection of theta, u, v, w with Omega = 0 (advection.py:134-182),
// the constant-K diffusion, and the uncoupled update from the time-t
// copies with w, phi' and mu' frozen (dycore.py:1602-1643).  Dry,
// periodic, msf = 1; the caller guarantees w == 0.  A test path of the
// reference: thread-per-point kernels over the union of the staggers.
// ---------------------------------------------------------------------------

// Stage preparation: tendencies zeroed, the total theta into th_pp, the
// coupled fluxes (c1h <mu> + c2h) u, v into u_pp, v_pp, Omega = 0 in ww_pp.
__global__ void k_adv_prep(Dom d)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    if (i < nx && j < ny && k < nz) {
        const int x = iM(k, j, i, ny, nx), c = j * nx + i;
        d.th_pp[x] = basev(d.thb, k, c, st, d.base3d) + d.thp[x];
        d.rth_t[x] = R(0.0);
    }
    if (j < ny && k < nz) {
        const int x = iU(k, j, i, ny, nx);
        d.u_pp[x] = ru_at(d, k, j, i);
        d.ru_t[x] = R(0.0);
    }
    if (i < nx && k < nz) {
        const int x = iV(k, j, i, ny, nx);
        d.v_pp[x] = rv_at(d, k, j, i);
        d.rv_t[x] = R(0.0);
    }
    if (i < nx && j < ny) {
        const int x = iM(k, j, i, ny, nx);
        d.ww_pp[x] = R(0.0);
        d.rw_t[x] = R(0.0);
        d.rph_t[x] = R(0.0);
        if (k == 0) d.rmu_t[j * nx + i] = R(0.0);
    }
}

// Constant-K Laplacian of field f at (k, j, i) of a (nlev, ny', nx') array
// with row pitch px: kh (f_xx + f_yy) + kv (f_z)_z (reference
// diffusion.cu:22-51), the same float32 expressions as the stage kernels'
// fused diffusion.  Vertical spacings rdzf/rdzc of the field's levels.
__device__ __forceinline__ real lap2(const real* __restrict__ f, int k, int j, int i,
                                     int nlev, int ny, int nx, int px,
                                     int plane, const Sc& s,
                                     const real* __restrict__ rdzf,
                                     const real* __restrict__ rdzc)
{
    const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j + 1 == ny) ? 0 : j + 1;
    const int im = (i == 0) ? nx - 1 : i - 1, ip = (i + 1 == nx) ? 0 : i + 1;
    const int r = k * plane;
    const real fc = f[r + j * px + i];
    real lap = s.kh * (s.dxi2 * ((f[r + j * px + ip] - R(2.0) * fc)
                                 + f[r + j * px + im])
                       + s.dyi2 * ((f[r + jp * px + i] - R(2.0) * fc)
                                   + f[r + jm * px + i]));
    const real fup = (k < nlev - 1)
        ? (f[r + plane + j * px + i] - fc) * rdzf[k] : R(0.0);
    const real fdn = (k > 0)
        ? (fc - f[r - plane + j * px + i]) * rdzf[k - 1] : R(0.0);
    return lap + (s.kv * (fup - fdn)) * rdzc[k];
}

// The diffusion tendencies coupled with C(mu) at each staggering (mass
// at the face for u, v), added to ru_t, rv_t, rw_t, rth_t.
__global__ void k_adv_diff(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const real c1 = k < nz ? d.c1h[k] : R(0.0), c2 = k < nz ? d.c2h[k] : R(0.0);
    if (j < ny && k < nz) {              // u faces, the duplicate included
        const int ic = (i >= nx) ? i - nx : i, ib = (ic == 0) ? nx - 1 : ic - 1;
        const real mux = R(0.5) * (mut_at(d, j * nx + ic) + mut_at(d, j * nx + ib));
        const real lap = lap2(d.u, k, j, ic, nz, ny, nx, nx + 1,
                              ny * (nx + 1), s, d.rdzf_h, d.rdzc_h);
        const int x = iU(k, j, i, ny, nx);
        d.ru_t[x] = d.ru_t[x] + (c1 * mux + c2) * lap;
    }
    if (i < nx && k < nz) {              // v faces, the duplicate row included
        const int jc = (j >= ny) ? j - ny : j, jb = (jc == 0) ? ny - 1 : jc - 1;
        const real muy = R(0.5) * (mut_at(d, jc * nx + i) + mut_at(d, jb * nx + i));
        const real lap = lap2(d.v, k, jc, i, nz, ny, nx, nx, (ny + 1) * nx, s,
                              d.rdzf_h, d.rdzc_h);
        const int x = iV(k, j, i, ny, nx);
        d.rv_t[x] = d.rv_t[x] + (c1 * muy + c2) * lap;
    }
    if (i < nx && j < ny) {
        const int c = j * nx + i;
        const real mu = mut_at(d, c);
        if (k < nz) {
            const real lap = lap2(d.thp, k, j, i, nz, ny, nx, nx, st, s,
                                  d.rdzf_h, d.rdzc_h);
            const int x = iM(k, j, i, ny, nx);
            d.rth_t[x] = d.rth_t[x] + (c1 * mu + c2) * lap;
        }
        // w: interior levels only; the surface and lid get zero
        real lap = R(0.0);
        if (k > 0 && k < nz)
            lap = lap2(d.w, k, j, i, nz + 1, ny, nx, nx, st, s, d.rdzf_w,
                       d.rdzc_w);
        const int x = iM(k, j, i, ny, nx);
        d.rw_t[x] = d.rw_t[x] + (d.c1f[k] * mu + d.c2f[k]) * lap;
    }
}

// Uncoupled update from the time-t copies (reference dycore.py:1602-1643):
// q = (C(mu_t) q_t + dt r_q) / C(mu), face masses for u, v.
__global__ void k_adv_advance(Dom d, real dt)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    auto m0 = [&](int c) { return d.mub2d[c] + d.mup0[c]; };
    auto m1 = [&](int c) { return d.mub2d[c] + d.mup[c]; };
    if (i < nx && j < ny && k < nz) {
        const int x = iM(k, j, i, ny, nx), c = j * nx + i;
        const real thb = basev(d.thb, k, c, st, d.base3d);
        d.thp[x] = (((d.c1h[k] * m0(c) + d.c2h[k]) * (thb + d.thp0[x])
                     + dt * d.rth_t[x]) / (d.c1h[k] * m1(c) + d.c2h[k]))
                   - thb;
    }
    if (j < ny && k < nz) {
        const int a = (i >= nx) ? i - nx : i, b = (a == 0) ? nx - 1 : a - 1;
        const real x0 = R(0.5) * (m0(j * nx + a) + m0(j * nx + b));
        const real x1 = R(0.5) * (m1(j * nx + a) + m1(j * nx + b));
        const int x = iU(k, j, i, ny, nx);
        d.u[x] = ((d.c1h[k] * x0 + d.c2h[k]) * d.u0[x] + dt * d.ru_t[x])
                 / (d.c1h[k] * x1 + d.c2h[k]);
    }
    if (i < nx && k < nz) {
        const int a = (j >= ny) ? j - ny : j, b = (a == 0) ? ny - 1 : a - 1;
        const real y0 = R(0.5) * (m0(a * nx + i) + m0(b * nx + i));
        const real y1 = R(0.5) * (m1(a * nx + i) + m1(b * nx + i));
        const int x = iV(k, j, i, ny, nx);
        d.v[x] = ((d.c1h[k] * y0 + d.c2h[k]) * d.v0[x] + dt * d.rv_t[x])
                 / (d.c1h[k] * y1 + d.c2h[k]);
    }
    if (i < nx && j < ny) {
        const int x = iM(k, j, i, ny, nx), c = j * nx + i;
        d.w[x] = ((d.c1f[k] * m0(c) + d.c2f[k]) * d.w0[x] + dt * d.rw_t[x])
                 / (d.c1f[k] * m1(c) + d.c2f[k]);
    }
}

extern "C" {

int32_t ARW_API(abi_version)(void) { return ARW_ABI_VERSION; }

const char* ARW_API(last_error)(void) { return g_err; }

int32_t ARW_API(calc_p_alpha)(const Dom* d, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    hipLaunchKernelGGL(k_eos<false>, col_grid(*d), dim3(256), 0,
                       (hipStream_t)stream, *d);
    return launched("k_eos");
}

int32_t ARW_API(stage_fluxes)(const Dom* d, const arw_params* prm,
                         real* ru, real* rv, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !ru || !rv) return fail(ARW_E_ARG, "null argument");
    if ((rc = check_prec(prm, d))) return rc;
    Sc s = make_sc(prm, prm->dt, 1);
    hipStream_t st = (hipStream_t)stream;
    if ((rc = launched(run_omega(*d, s, 0, st), "k_omega", -1, st))) return rc;
    hipLaunchKernelGGL(k_faces, union_grid(*d, d->nz), dim3(64, 4), 0, st,
                       *d, ru, rv);
    return launched("k_faces");
}

static int32_t flux_div(int which, const real* q, const real* ru,
                        const real* rv, const real* rw, real* tend,
                        const real* rdz, const real* fnm, const real* fnp,
                        real dx_inv, real dy_inv, int32_t nz, int32_t ny,
                        int32_t nx, void* stream)
{
    if (!q || !ru || !rv || !rw || !tend || !rdz || !fnm || !fnp)
        return fail(ARW_E_ARG, "null argument");
    if (nz < 1 || ny < 1 || nx < 1) return fail(ARW_E_ARG, "bad extents");
    if ((int64_t)(nz + 1) * (ny + 1) * (nx + 1) >= ((int64_t)1 << 31))
        return fail(ARW_E_ARG, "extents exceed 32-bit indexing");
    dim3 grid((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz + 1);
    hipStream_t st = (hipStream_t)stream;
    switch (which) {
    case 0:
        hipLaunchKernelGGL(k_flux_div<0>, grid, dim3(64, 4), 0, st, q, ru, rv,
                           rw, tend, rdz, fnm, fnp, dx_inv, dy_inv, nz, ny, nx);
        break;
    case 1:
        hipLaunchKernelGGL(k_flux_div<1>, grid, dim3(64, 4), 0, st, q, ru, rv,
                           rw, tend, rdz, fnm, fnp, dx_inv, dy_inv, nz, ny, nx);
        break;
    case 2:
        hipLaunchKernelGGL(k_flux_div<2>, grid, dim3(64, 4), 0, st, q, ru, rv,
                           rw, tend, rdz, fnm, fnp, dx_inv, dy_inv, nz, ny, nx);
        break;
    default:
        hipLaunchKernelGGL(k_flux_div<3>, grid, dim3(64, 4), 0, st, q, ru, rv,
                           rw, tend, rdz, fnm, fnp, dx_inv, dy_inv, nz, ny, nx);
    }
    return launched("k_flux_div");
}

int32_t ARW_API(flux_div_scalar)(const real* q, const real* ru, const real* rv,
                            const real* rw, real* tend, const real* rdnw,
                            const real* fnm, const real* fnp, real dx_inv,
                            real dy_inv, int32_t nz, int32_t ny, int32_t nx,
                            void* stream)
{
    return flux_div(0, q, ru, rv, rw, tend, rdnw, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t ARW_API(flux_div_u)(const real* u, const real* ru, const real* rv,
                       const real* rw, real* tend, const real* rdnw,
                       const real* fnm, const real* fnp, real dx_inv,
                       real dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void* stream)
{
    return flux_div(1, u, ru, rv, rw, tend, rdnw, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t ARW_API(flux_div_v)(const real* v, const real* ru, const real* rv,
                       const real* rw, real* tend, const real* rdnw,
                       const real* fnm, const real* fnp, real dx_inv,
                       real dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void* stream)
{
    return flux_div(2, v, ru, rv, rw, tend, rdnw, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t ARW_API(flux_div_w)(const real* w, const real* ru, const real* rv,
                       const real* rw, real* tend, const real* rdn,
                       const real* fnm, const real* fnp, real dx_inv,
                       real dy_inv, int32_t nz, int32_t ny, int32_t nx,
                       void* stream)
{
    return flux_div(3, w, ru, rv, rw, tend, rdn, fnm, fnp, dx_inv, dy_inv,
                    nz, ny, nx, stream);
}

int32_t ARW_API(stage_prepare)(const Dom* d, const arw_params* prm,
                          double dtau, int32_t do_eos, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if ((rc = check_params(prm, d))) return rc;
    if ((rc = refuse_cq(prm))) return rc;
    // the held 6th-order tendencies are formed and added by the whole step
    // only; a piecewise caller would silently miss them
    if (prm->diff_6th_opt != 0 || prm->km_opt == 4)
        return fail(ARW_E_UNSUPPORTED, "stage_prepare: diff_6th_opt and "
                                       "km_opt = 4 are applied by arw_step only");
    Sc s = make_sc(prm, dtau, 1);
    hipStream_t st = (hipStream_t)stream;
    if ((rc = stage_prepare(*d, s, do_eos, st))) return rc;
    // the reference initialises p_pp_old = p_pp (dycore.cu:1083-1084)
    size_t bytes = (size_t)d->nz * d->ny * d->nx * sizeof(real);
    if (hipMemcpyAsync(d->p_pp_old, d->p_pp, bytes, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return fail(ARW_E_LAUNCH, "p_pp_old copy failed");
    return ARW_OK;
}

int32_t ARW_API(calc_coefs)(const Dom* d, const arw_params* prm,
                       double dtau, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm) return fail(ARW_E_ARG, "null params");
    if ((rc = check_prec(prm, d))) return rc;
    if ((rc = refuse_cq(prm))) return rc;
    // The implicit-w factors are formed inside every substep from the
    // stage reference p/alt; what persists across substeps is the
    // reference mass divergence (and Omega), formed here.
    Sc s = make_sc(prm, dtau, 1);
    hipStream_t st = (hipStream_t)stream;
    return launched(run_omega(*d, s, 0, st), "k_omega", -1, st);
}

int32_t ARW_API(acoustic_substep)(const Dom* d, const arw_params* prm,
                             double dtau, int32_t first, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm) return fail(ARW_E_ARG, "null params");
    if ((rc = check_prec(prm, d))) return rc;
    if ((rc = refuse_cq(prm))) return rc;
    Sc s = make_sc(prm, dtau, first);
    s.use_pold = 1;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = launched(run_uv(*d, s, st), "k_uv", -1, st))) return rc;
    // the reference keeps p_pp_old = p'' before the column update
    size_t bytes = (size_t)d->nz * d->ny * d->nx * sizeof(real);
    if (hipMemcpyAsync(d->p_pp_old, d->p_pp, bytes, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return fail(ARW_E_LAUNCH, "p_pp_old copy failed");
    return launched(run_column(*d, s, d->p_pp, 1, 1, st), "k_column", -1, st);
}

int32_t ARW_API(small_step_finish)(const Dom* d, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    FinishDst o = {d->u, d->v, d->w, d->thp, d->php, d->mup_next};
    if ((rc = launched(run_finish(*d, o, false, st), "k_finish", -1, st))) return rc;
    if (hipMemcpyAsync(d->mup, d->mup_next,
                       (size_t)d->ny * d->nx * sizeof(real),
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(ARW_E_LAUNCH, "mup copy failed");
    return ARW_OK;
}

int32_t ARW_API(set_w_surface)(const Dom* d, const arw_params* prm,
                          void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm) return fail(ARW_E_ARG, "null params");
    Sc s = make_sc(prm, prm->dt, 1);
    hipLaunchKernelGGL(k_set_w_surface<One>, col_grid(*d), dim3(256), 0,
                       (hipStream_t)stream, *d, s);
    return launched("k_set_w_surface");
}

// Whole step.  Instead of copying the time-t fields, the stage buffers
// rotate: with A = the caller's prognostic arrays (time t) and B = the
// *0 arrays, stage 1 reads A and finishes into B, stage 2 finishes B in
// place, stage 3 reads B and finishes into A; the time-t fields the stage
// initialisations need stay in A throughout.  mu' rotates A -> B ->
// mup_next -> A because the finish reads neighbouring mu' while it writes.
// After the step the *0 arrays hold the stage-2 estimate (they are
// rebuilt every step, like the reference's time-t copies).
// ck (optional) receives the stage checkpoints of the scalar transport:
// the sumflux averages accumulate straight into its per-stage buffers.
static int32_t step_impl(const Dom* d, const arw_params* prm, Ckpt* ck,
                         void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if ((rc = check_params(prm, d))) return rc;
    const bool prec = (!ARW_F64) && prm->precision_form != 0;
    if (prm->diff_6th_opt != 0) {
        if (prm->diff_6th_opt != 1 && prm->diff_6th_opt != 2)
            return fail(ARW_E_ARG, "diff_6th_opt must be 0, 1 or 2");
        if (!d->d6_u || !d->d6_v || !d->d6_w || !d->d6_t)
            return fail(ARW_E_ARG, "diff_6th_opt needs the d6_* buffers");
        if (prec) return fail(ARW_E_ARG, "diff_6th_opt: plain form only");
        if (d->qv && !d->held_q)
            return fail(ARW_E_ARG, "diff_6th_opt on a moist state needs "
                                   "held_q");
    }
    if (ck) {
        if (!ck->mup_t) return fail(ARW_E_ARG, "checkpoint: null mup_t");
        for (int s = 0; s < 3; ++s)
            if (!ck->ru_m[s] || !ck->rv_m[s] || !ck->ww_m[s] || !ck->mup_new[s])
                return fail(ARW_E_ARG, "checkpoint: null stage buffer");
    }
    hipStream_t st = (hipStream_t)stream;
    prof_mark(KID_COUNT, st);                    // opening event
    const int ns = prm->time_step_sound;
    const int nsub[3] = {1, ns / 2 > 1 ? ns / 2 : 1, ns};
    const double dtaus[3] = {prm->dt / 3.0, prm->dt / ns, prm->dt / ns};
    real* A[6] = {d->u, d->v, d->w, d->thp, d->php, d->mup};
    real* B[6] = {d->u0, d->v0, d->w0, d->thp0, d->php0, d->mup0};
    real* mup_c = d->mup_next;
    const bool moist = d->qv != nullptr;
    // every transported species: qv, qc, qr, then the scheme's extras
    const int nq = 3 + (moist ? d->n_extra : 0);
    real* QA[3 + ARW_MAX_EXTRA] = {d->qv, d->qc, d->qr};
    real* QB[3 + ARW_MAX_EXTRA] = {d->qv0, d->qc0, d->qr0};
    real* QC[3 + ARW_MAX_EXTRA] = {d->qsv, d->qsc, d->qsr};
    for (int e = 3; e < nq; ++e) {
        QA[e] = d->qx[e - 3];
        QB[e] = d->qx0[e - 3];
        QC[e] = d->qxs[e - 3];
    }
    const bool sum = moist || ck;                // substep flux averages
    if (ck && hipMemcpyAsync(ck->mup_t, A[5], (size_t)d->ny * d->nx * RB,
                             hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(ARW_E_LAUNCH, "checkpoint mu' copy failed");
    const bool smag = prm->km_opt == 4;
    if (smag) {
        // the Smagorinsky density 1/alt is the time-t state's: refresh the
        // EOS first (reference dycore.py:2353-2354)
        if ((rc = launched(run_eos(*d, st), "k_eos", -1, st))) return rc;
        if ((rc = launched(run_smag(*d, prm, d->sm_buf, d->d6_u, d->d6_v,
                                    d->d6_w, d->d6_t, QA, d->held_q,
                                    moist ? nq : 0, st),
                           "smag_held", KID_SMAG, st)))
            return rc;
    }
    const bool d6 = smag || prm->diff_6th_opt > 0;
    if (prm->diff_6th_opt > 0
        && (rc = held_diff6(*d, prm, st, smag, QA, moist ? nq : 0)))
        return rc;
    // the species' held tendencies: Smagorinsky's for every species, the
    // filter's for those it acts on; null where nothing is held
    real* fixed[3 + ARW_MAX_EXTRA] = {};
    for (int e = 0; moist && e < nq; ++e)
        if (smag || (prm->diff_6th_opt > 0 && ((prm->diff6_species >> e) & 1u)))
            fixed[e] = d->held_q + (size_t)e * d->nz * d->ny * d->nx;
    for (int stage = 0; stage < 3; ++stage) {
        Dom dd = *d;
        // time-t fields: always A
        dd.u0 = A[0]; dd.v0 = A[1]; dd.w0 = A[2]; dd.thp0 = A[3];
        dd.php0 = A[4]; dd.mup0 = A[5];
        // stage reference t*
        real* const* T = (stage == 0) ? A : B;
        dd.u = T[0]; dd.v = T[1]; dd.w = T[2]; dd.thp = T[3]; dd.php = T[4];
        dd.mup = (stage == 0) ? A[5] : (stage == 1 ? B[5] : mup_c);
        FinishDst o;
        real* const* D = (stage == 2) ? A : B;
        o.u = D[0]; o.v = D[1]; o.w = D[2]; o.thp = D[3]; o.php = D[4];
        o.mup = (stage == 0) ? B[5] : (stage == 1 ? mup_c : A[5]);

        if (moist) {
            // scalars of the stage state: time t, then the stage buffers
            real* const* Q = (stage == 0) ? QA : (stage == 1 ? QB : QC);
            dd.qv = Q[0]; dd.qc = Q[1]; dd.qr = Q[2];
            for (int e = 3; e < nq; ++e) dd.qx[e - 3] = Q[e];
        }
        if (ck) {
            dd.ru_m = ck->ru_m[stage]; dd.rv_m = ck->rv_m[stage];
            dd.ww_m = ck->ww_m[stage];
            ck->dt[stage] = nsub[stage] * dtaus[stage];
        }
        Sc s = make_sc(prm, dtaus[stage], 1);
        if (s.cq) {
            // moist_cq: the stage's cq from its scalars, before the slow
            // pressure gradient reads it
            hipLaunchKernelGGL(k_calc_cq, union_grid(dd, dd.nz + 1),
                               dim3(64, 4), 0, st, dd, prm->moist_cq);
            if ((rc = launched(hipGetLastError(), "k_calc_cq", -1, st)))
                return rc;
        }
        // the held tendencies join the stage's as the stage kernels store
        // them, after the constant-K diffusion (reference dycore.py:
        // 2389-2390, add_fixed_dry_tendencies :1422-1430)
        s.held = d6;
        if ((rc = stage_prepare(dd, s, stage == 0, st))) return rc;
        Dom ds = dd;
        for (int n = 0; n < nsub[stage]; ++n) {
            Sc sn = make_sc(prm, dtaus[stage], n == 0);
            sn.use_pold = (n > 0);
            sn.gam_ready = (n > 0);             // stored by the first
            // WRF sumflux (reference acoustic.cu:1437-1468) inside the
            // substep pass: the first substep starts the sums
            sn.sum = sum ? (n == 0 ? 1 : 2) : 0;
            const bool last = (n + 1 == nsub[stage]);
            // new p'' goes to the buffer holding the dead p''(n-1); the two
            // names swap afterwards
            if ((rc = substep(ds, sn, ds.p_pp_old, ds.base3d || last, sum,
                              st)))
                return rc;
            real* t = ds.p_pp;
            ds.p_pp = ds.p_pp_old;
            ds.p_pp_old = t;
        }
        if (sum) {
            // WRF sumflux finish, then the scalar stage update -- both
            // before the finish, which on the last stage overwrites the
            // time-t mass the update divides by
            hipLaunchKernelGGL(k_sumflux_finish, union_grid(*d, d->nz + 1),
                               dim3(64, 4), 0, st, dd, (real)nsub[stage]);
            if ((rc = launched(hipGetLastError(), "k_sumflux_finish", KID_SCALARS, st)))
                return rc;
        }
        if (ck) {
            hipLaunchKernelGGL(k_mup_new, col_grid(*d), dim3(256), 0, st, dd,
                               ck->mup_new[stage]);
            if ((rc = launched(hipGetLastError(), "k_mup_new", KID_SCALARS, st)))
                return rc;
        }
        if (moist) {
            real* const* Q = (stage == 0) ? QA : (stage == 1 ? QB : QC);
            real* const* Dst = (stage == 0) ? QB : (stage == 1 ? QC : QA);
            // QB (stage 1's output) is dead during the final stage
            if ((rc = scalar_stage(dd, prm, Q, QA, Dst, A[5],
                                   nsub[stage] * dtaus[stage], stage == 2, st,
                                   nq, QB, fixed)))
                return rc;
        }
        // the final stage removes the heating its substeps integrated
        const double hdiab_dt = (stage == 2 && d->h_diabatic)
                                    ? nsub[stage] * dtaus[stage] : 0.0;
        if ((rc = launched(run_finish(dd, o, prec, st, hdiab_dt), "k_finish",
                           KID_FINISH, st)))
            return rc;
        if (moist) {
            // moist EOS of the new stage state (needs its updated qv)
            Dom de = dd;
            de.thp = o.thp; de.php = o.php; de.mup = o.mup;
            de.qv = (stage == 0) ? QB[0] : (stage == 1 ? QC[0] : QA[0]);
            if (prec)
                hipLaunchKernelGGL(k_eos<true>, col_grid(*d), dim3(256), 0, st, de);
            else
                hipLaunchKernelGGL(k_eos<false>, col_grid(*d), dim3(256), 0, st, de);
            if ((rc = launched(hipGetLastError(), "k_eos", KID_SCALARS, st)))
                return rc;
        }
    }
    Sc s = make_sc(prm, prm->dt, 1);
    hipLaunchKernelGGL(k_set_w_surface<One>, col_grid(*d), dim3(256), 0, st, *d,
                       s);
    return launched(hipGetLastError(), "k_set_w_surface", KID_SET_W_SURFACE,
                    st);
}

// The advection-only step (reference dycore.py:2318-2343); see
// k_adv_prep.  Uses u_pp, v_pp, ww_pp, th_pp as its flux and total-theta
// work arrays.
int32_t ARW_API(step_advective)(const Dom* d, const arw_params* prm,
                                void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if ((rc = check_params(prm, d))) return rc;
    if (d->qv)
        return fail(ARW_E_UNSUPPORTED, "step_advective: dry states only");
    if ((!ARW_F64 && prm->precision_form) || prm->diff_6th_opt
        || prm->km_opt == 4)
        return fail(ARW_E_UNSUPPORTED, "step_advective: plain form, no "
                                       "6th-order filter, no km_opt = 4");
    hipStream_t st = (hipStream_t)stream;
    const size_t m3 = (size_t)d->nz * d->ny * d->nx * sizeof(real);
    const size_t u3 = (size_t)d->nz * d->ny * (d->nx + 1) * sizeof(real);
    const size_t v3 = (size_t)d->nz * (d->ny + 1) * d->nx * sizeof(real);
    const size_t w3 = (size_t)(d->nz + 1) * d->ny * d->nx * sizeof(real);
    const size_t m2 = (size_t)d->ny * d->nx * sizeof(real);
    const struct { real* dst; const real* src; size_t n; } cp[] = {
        {d->u0, d->u, u3}, {d->v0, d->v, v3}, {d->w0, d->w, w3},
        {d->thp0, d->thp, m3}, {d->php0, d->php, w3}, {d->mup0, d->mup, m2}};
    for (const auto& c : cp)
        if (hipMemcpyAsync(c.dst, c.src, c.n, hipMemcpyDeviceToDevice, st)
            != hipSuccess)
            return fail(ARW_E_LAUNCH, "step_advective: time-t copy failed");
    const Sc s = make_sc(prm, prm->dt, 1);
    const bool diff = s.diff;
    const dim3 g = union_grid(*d, d->nz + 1), b(64, 4);
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const double lens[3] = {prm->dt / 3.0, prm->dt / 2.0, prm->dt};
    const int nz = d->nz, ny = d->ny, nx = d->nx;
    for (int stage = 0; stage < 3; ++stage) {
        if ((rc = launched(run_eos(*d, st), "k_eos", -1, st))) return rc;
        hipLaunchKernelGGL(k_adv_prep, g, b, 0, st, *d);
        if ((rc = launched("k_adv_prep"))) return rc;
        if ((rc = flux_div(0, d->th_pp, d->u_pp, d->v_pp, d->ww_pp, d->rth_t,
                           d->rdnw, d->fnm, d->fnp, rdx, rdy, nz, ny, nx, st)))
            return rc;
        if ((rc = flux_div(1, d->u, d->u_pp, d->v_pp, d->ww_pp, d->ru_t,
                           d->rdnw, d->fnm, d->fnp, rdx, rdy, nz, ny, nx, st)))
            return rc;
        if ((rc = flux_div(2, d->v, d->u_pp, d->v_pp, d->ww_pp, d->rv_t,
                           d->rdnw, d->fnm, d->fnp, rdx, rdy, nz, ny, nx, st)))
            return rc;
        if ((rc = flux_div(3, d->w, d->u_pp, d->v_pp, d->ww_pp, d->rw_t,
                           d->rdn, d->fnm, d->fnp, rdx, rdy, nz, ny, nx, st)))
            return rc;
        if (diff) {
            hipLaunchKernelGGL(k_adv_diff, g, b, 0, st, *d, s);
            if ((rc = launched("k_adv_diff"))) return rc;
        }
        hipLaunchKernelGGL(k_adv_advance, g, b, 0, st, *d, (real)lens[stage]);
        if ((rc = launched("k_adv_advance"))) return rc;
    }
    hipLaunchKernelGGL(k_set_w_surface<One>, col_grid(*d), dim3(256), 0, st,
                       *d, s);
    return launched("k_set_w_surface");
}

int32_t ARW_API(step)(const Dom* d, const arw_params* prm, void* stream)
{
    return step_impl(d, prm, nullptr, stream);
}

int32_t ARW_API(step_ckpt)(const Dom* d, const arw_params* prm, Ckpt* ck,
                           void* stream)
{
    if (!ck) return fail(ARW_E_ARG, "null checkpoint");
    return step_impl(d, prm, ck, stream);
}

// ---------------------------------------------------------------------------
// The step with library-held scratch, for callers that cannot allocate.
// Its four scratch arrays (Omega, the column solve's factors, the
// reference mass divergence, the next stage's mu') live in one file-scope
// device array of kStepScratchBytes per GPU, held until the process ends.
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU race.  Each array is written before it is read within a step, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr size_t kStepScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStepScratchBytes];

constexpr int kMaxDevices = 64;
static char* g_scratch_dev[kMaxDevices];          // per device, set once
static std::mutex g_scratch_mu;

// This device's address of `scratch` (the host-side name is only a
// stand-in for the symbol).
static int32_t step_scratch(char** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(ARW_E_LAUNCH, "step_noalloc: no current device");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            snprintf(g_err, sizeof(g_err),
                     "step_noalloc: hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
            return ARW_E_LAUNCH;
        }
        g_scratch_dev[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_dev[dev];
    return ARW_OK;
}

int32_t ARW_API(step_noalloc)(const Dom* d, const arw_params* prm,
                              void* stream)
{
    if (!d) return fail(ARW_E_ARG, "null domain");
    if (!prm) return fail(ARW_E_ARG, "null params");
    if (d->qv || (!ARW_F64 && prm->precision_form) || prm->diff_6th_opt
        || prm->km_opt == 4)
        return fail(ARW_E_UNSUPPORTED,
                    "step_noalloc: dry plain-form steps without the "
                    "6th-order filter or km_opt = 4 only");
    if (d->nz < 3 || d->ny < 1 || d->nx < 1)
        return fail(ARW_E_ARG, "need nz >= 3, ny >= 1, nx >= 1");
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t col = (size_t)d->ny * d->nx * sizeof(real);
    const size_t lev = (size_t)(d->nz + 1) * col;
    const size_t need = 2 * al(lev) + 2 * al(col);
    if (need > kStepScratchBytes) {
        snprintf(g_err, sizeof(g_err),
                 "step_noalloc: this domain needs %zu bytes of step scratch; "
                 "the library holds %zu per GPU",
                 need, kStepScratchBytes);
        return ARW_E_ARG;
    }
    char* base = nullptr;
    if (int32_t rc = step_scratch(&base)) return rc;
    Dom dd = *d;
    dd.ww = reinterpret_cast<real*>(base);
    dd.gam = reinterpret_cast<real*>(base + al(lev));
    dd.dmdt_rf = reinterpret_cast<real*>(base + 2 * al(lev));
    dd.mup_next = reinterpret_cast<real*>(base + 2 * al(lev) + al(col));
    return step_impl(&dd, prm, nullptr, stream);
}

static int32_t check_tracer(const Dom* d, const arw_params* prm,
                            const Ckpt* ck, const void* a, const void* b,
                            const void* work)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !(prm->dx > 0.0) || !(prm->dy > 0.0))
        return fail(ARW_E_ARG, "params: dx, dy must be positive");
    if (!ck || !ck->mup_t) return fail(ARW_E_ARG, "null checkpoint");
    for (int s = 0; s < 3; ++s)
        if (!ck->ru_m[s] || !ck->rv_m[s] || !ck->ww_m[s] || !ck->mup_new[s])
            return fail(ARW_E_ARG, "checkpoint: null stage buffer");
    if (!a || !b || !work) return fail(ARW_E_ARG, "null scalar or work array");
    return ARW_OK;
}

int32_t ARW_API(tracer_step)(const Dom* d, const arw_params* prm,
                             const Ckpt* ck, const real* q_in, real* q_out,
                             real* work, int32_t clamp, void* stream)
{
    int32_t rc = check_tracer(d, prm, ck, q_in, q_out, work);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)d->nz * d->ny * d->nx;
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const dim3 cells((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
    const real* cur[3] = {q_in, work, work + n};
    real* dst[3] = {work, work + n, q_out};
    for (int s = 0; s < 3; ++s) {
        hipLaunchKernelGGL(k_tracer_stage, cells, dim3(64, 4), 0, st, *d,
                           cur[s], q_in, dst[s], ck->ru_m[s], ck->rv_m[s],
                           ck->ww_m[s], ck->mup_t, ck->mup_new[s],
                           (real)ck->dt[s], (clamp && s == 2) ? 1 : 0, rdx,
                           rdy);
        if ((rc = launched("k_tracer_stage"))) return rc;
    }
    return ARW_OK;
}

// lam_out = J^T lam_in for J = d tracer_step / d q_in (no active clamp).
// Stage s maps g -> m = g / C(mu_s), gq += C(mu_t) m, g = dt_s D_s^T m,
// from the last stage back; lam_out = gq + g.
int32_t ARW_API(tracer_adjoint)(const Dom* d, const arw_params* prm,
                                const Ckpt* ck, const real* lam_in,
                                real* lam_out, real* work, void* stream)
{
    int32_t rc = check_tracer(d, prm, ck, lam_in, lam_out, work);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)d->nz * d->ny * d->nx;
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const dim3 cells((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
    real* m = work;
    real* gq = work + n;
    real* g = work + 2 * n;
    for (int s = 2; s >= 0; --s) {
        hipLaunchKernelGGL(k_tracer_adj_scale, cells, dim3(64, 4), 0, st, *d,
                           s == 2 ? lam_in : g, m, gq, ck->mup_t,
                           ck->mup_new[s], s == 2 ? 1 : 0);
        if ((rc = launched("k_tracer_adj_scale"))) return rc;
        hipLaunchKernelGGL(k_tracer_adj_div, cells, dim3(64, 4), 0, st, *d, m,
                           g, ck->ru_m[s], ck->rv_m[s], ck->ww_m[s],
                           (real)ck->dt[s], rdx, rdy);
        if ((rc = launched("k_tracer_adj_div"))) return rc;
    }
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_add3, dim3(blocks), dim3(256), 0, st, lam_out, gq, g,
                       (int)n);
    return launched("k_add3");
}

// One stage of the unlimited scalar update with explicit fluxes and
// masses (the arithmetic of k_scalar_update; reference moist.py:385-398):
// q_new = (C(mu_t) q0 + dt flux_div(q; ru, rv, ww)) / C(mu_new), clamped
// at zero when clamp.
int32_t ARW_API(scalar_stage)(const Dom* d, const arw_params* prm,
                              double dt_stage, int32_t clamp, const real* q,
                              const real* q0, const real* ru, const real* rv,
                              const real* ww, const real* mup_t,
                              const real* mup_new, real* q_new, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !q || !q0 || !ru || !rv || !ww || !mup_t || !mup_new || !q_new)
        return fail(ARW_E_ARG, "null argument");
    if (q_new == q) return fail(ARW_E_ARG, "q_new must not alias q");
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const dim3 cells((d->nx + 63) / 64, (d->ny + 3) / 4, d->nz);
    hipLaunchKernelGGL(k_tracer_stage, cells, dim3(64, 4), 0,
                       (hipStream_t)stream, *d, q, q0, q_new, ru, rv, ww,
                       mup_t, mup_new, (real)dt_stage, clamp ? 1 : 0, rdx,
                       rdy);
    return launched("k_tracer_stage");
}

// Its transpose for lam = adjoint of q_new: g_q is written; g_q0, the flux
// gradients g_ru, g_rv, g_ww and the column-mass gradients g_mup_t,
// g_mup_new (mu' of time t and after the stage) are added to.  The clamp
// is differentiated as 1 where q_new > 0, 0 elsewhere.  work: 3 nz*ny*nx.
int32_t ARW_API(scalar_stage_adjoint)(
    const Dom* d, const arw_params* prm, double dt_stage, int32_t clamp,
    const real* q, const real* q0, const real* ru, const real* rv,
    const real* ww, const real* mup_t, const real* mup_new, const real* lam,
    real* g_q, real* g_q0, real* g_ru, real* g_rv, real* g_ww, real* g_mup_t,
    real* g_mup_new, real* work, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !q || !q0 || !ru || !rv || !ww || !mup_t || !mup_new || !lam
        || !g_q || !g_q0 || !g_ru || !g_rv || !g_ww || !g_mup_t || !g_mup_new
        || !work)
        return fail(ARW_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int nz = d->nz, ny = d->ny, nx = d->nx;
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const dim3 cells((nx + 63) / 64, (ny + 3) / 4, nz);
    hipLaunchKernelGGL(k_ss_adj_pts, cells, dim3(64, 4), 0, st, *d, q, q0, ru,
                       rv, ww, mup_t, mup_new, lam, g_q0, work,
                       (real)dt_stage, clamp ? 1 : 0, rdx, rdy);
    if ((rc = launched("k_ss_adj_pts"))) return rc;
    hipLaunchKernelGGL(k_ss_adj_cols, dim3((ny * nx + 63) / 64), dim3(64, KG),
                       0, st, *d, work, g_mup_t, g_mup_new);
    if ((rc = launched("k_ss_adj_cols"))) return rc;
    if (hipMemsetAsync(g_q, 0, (size_t)nz * ny * nx * RB, st) != hipSuccess)
        return fail(ARW_E_LAUNCH, "scalar adjoint: clearing g_q");
    return lib_flux_div_adjoint(0, q, ru, rv, ww, work, g_q, g_ru, g_rv, g_ww,
                                d->rdnw, d->fnm, d->fnp, rdx, rdy, nz, ny, nx,
                                st);
}

// The final stage's update of one scalar with WRF's positive-definite
// limiter and explicit fluxes and masses (the arithmetic of the step's
// moist_adv_opt = 1 path; reference pd_advection.cu, moist.py:385-398,
// :493-494).  work: 26 (nz+1)(ny+1)(nx+1) elements.
int32_t ARW_API(pd_stage)(const Dom* d, const arw_params* prm,
                          double dt_stage, const real* q, const real* q0,
                          const real* ru, const real* rv, const real* ww,
                          const real* mup_t, const real* mup_new,
                          real* q_new, real* work, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !q || !q0 || !ru || !rv || !ww || !mup_t || !mup_new
        || !q_new || !work)
        return fail(ARW_E_ARG, "null argument");
    if (q_new == q || q_new == q0)
        return fail(ARW_E_ARG, "q_new must not alias q or q0");
    return pd_forward(d, prm, dt_stage, q, q0, ru, rv, ww, mup_t, mup_new,
                      q_new, pd_layout(*d, work), (hipStream_t)stream);
}

// Its transpose for lam = adjoint of q_new: g_q is written; g_q0, the flux
// gradients and the column-mass gradients g_mup_t, g_mup_new are added
// to.  Piecewise: each branch (clamp, donor choice, factor clamps, CFL
// caps) is differentiated as taken.  work: 26 (nz+1)(ny+1)(nx+1).
int32_t ARW_API(pd_stage_adjoint)(
    const Dom* d, const arw_params* prm, double dt_stage, const real* q,
    const real* q0, const real* ru, const real* rv, const real* ww,
    const real* mup_t, const real* mup_new, const real* lam, real* g_q,
    real* g_q0, real* g_ru, real* g_rv, real* g_ww, real* g_mup_t,
    real* g_mup_new, real* work, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!prm || !q || !q0 || !ru || !rv || !ww || !mup_t || !mup_new || !lam
        || !g_q || !g_q0 || !g_ru || !g_rv || !g_ww || !g_mup_t || !g_mup_new
        || !work)
        return fail(ARW_E_ARG, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int nz = d->nz, ny = d->ny, nx = d->nx;
    const PdWork w = pd_layout(*d, work);
    if ((rc = pd_forward(d, prm, dt_stage, q, q0, ru, rv, ww, mup_t, mup_new,
                         w.qn, w, st)))
        return rc;
    const real dt = (real)dt_stage;
    const real rdx = (real)(1.0 / prm->dx), rdy = (real)(1.0 / prm->dy);
    const dim3 cells((nx + 63) / 64, (ny + 3) / 4, nz);
    hipLaunchKernelGGL(k_pda_update, cells, dim3(64, 4), 0, st, *d, w, q0,
                       mup_t, mup_new, lam, g_q0, dt);
    if ((rc = launched("k_pda_update"))) return rc;
    hipLaunchKernelGGL(k_pda_scale, cells, dim3(64, 4), 0, st, *d, w, q0,
                       mup_t, g_q0, rdx, rdy, dt);
    if ((rc = launched("k_pda_scale"))) return rc;
    hipLaunchKernelGGL(k_pda_faces, union_grid(*d, nz + 1), dim3(64, 4), 0,
                       st, *d, w, q, q0, ru, rv, ww, mup_new, g_ru, g_rv,
                       g_ww, (real)prm->dx, (real)prm->dy, rdx, rdy, dt);
    if ((rc = launched("k_pda_faces"))) return rc;
    hipLaunchKernelGGL(k_pda_gather, cells, dim3(64, 4), 0, st, *d, w, ru,
                       rv, ww, g_q, g_q0);
    if ((rc = launched("k_pda_gather"))) return rc;
    hipLaunchKernelGGL(k_ss_adj_cols, dim3((ny * nx + 63) / 64), dim3(64, KG),
                       0, st, *d, w.tg, g_mup_t, g_mup_new);
    return launched("k_ss_adj_cols");
}

int32_t ARW_API(ensemble_create)(const Dom* members, int32_t n,
                                 const arw_params* prm, const int32_t* order,
                                 int32_t groups, int32_t col_tile,
                                 ARW_API(ensemble)** out)
{
    if (!out) return fail(ARW_E_ARG, "null output handle");
    *out = nullptr;
    if (prm && (prm->diff_6th_opt != 0 || prm->km_opt == 4))
        return fail(ARW_E_ARG, "ensemble: diff_6th_opt and km_opt = 4 are "
                               "not supported");
    if (!members || n < 1 || n > 65536)
        return fail(ARW_E_ARG, "need 1 <= n <= 65536 members");
    if (groups < 1 || groups > n)
        return fail(ARW_E_ARG, "need 1 <= groups <= n");
    if (col_tile != 64 && col_tile != 128 && col_tile != 256)
        return fail(ARW_E_ARG, "col_tile must be 64, 128 or 256");
    int32_t rc;
    for (int m = 0; m < n; ++m) {
        if ((rc = check_domain(&members[m]))) return rc;
        if ((rc = check_params(prm, &members[m]))) return rc;
        if (members[m].qv)
            return fail(ARW_E_ARG, "ensemble members must be dry");
        if ((members[m].base3d != 0) != (members[0].base3d != 0))
            return fail(ARW_E_ARG, "ensemble members must share base3d");
    }
    std::vector<int> ord(n);
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; ++i) {
        ord[i] = order ? order[i] : i;
        if (ord[i] < 0 || ord[i] >= n || seen[ord[i]])
            return fail(ARW_E_ARG, "order is not a permutation of 0..n-1");
        seen[ord[i]] = 1;
    }
    // contiguous groups of about equal work (mass points)
    double total_work = 0.0;
    for (int m = 0; m < n; ++m)
        total_work += (double)members[m].nz * members[m].ny * members[m].nx;
    std::vector<std::vector<int>> gm(groups);
    {
        double acc = 0.0;
        int g = 0;
        for (int i = 0; i < n; ++i) {
            // next group once this one holds its share of the work, or
            // when the members left are just enough for one per later group
            if (!gm[g].empty() && g + 1 < groups
                && (acc >= total_work * (g + 1) / groups
                    || n - i <= groups - 1 - g))
                ++g;
            const int m = ord[i];
            gm[g].push_back(m);
            acc += (double)members[m].nz * members[m].ny * members[m].nx;
        }
    }
    // host image of every group's views and tables, then one device copy
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    struct Off { size_t doms, fin, start[BK_COUNT], owner[BK_COUNT]; int total[BK_COUNT]; };
    std::vector<Off> off(groups);
    size_t bytes = 0;
    for (int g = 0; g < groups; ++g) {
        const size_t ng = gm[g].size();
        off[g].doms = bytes; bytes = align(bytes + ENS_ROLES * ng * sizeof(Dom));
        off[g].fin = bytes; bytes = align(bytes + 3 * ng * sizeof(FinishDst));
        for (int k = 0; k < BK_COUNT; ++k) {
            long long t = 0;
            for (int m : gm[g]) t += batch_blocks(members[m], k, col_tile);
            if (t >= (1LL << 30)) return fail(ARW_E_ARG, "ensemble too large for one launch");
            off[g].total[k] = (int)t;
            off[g].start[k] = bytes; bytes = align(bytes + (ng + 1) * sizeof(int));
            off[g].owner[k] = bytes; bytes = align(bytes + (size_t)t * sizeof(int));
        }
    }
    std::vector<unsigned char> img(bytes, 0);
    for (int g = 0; g < groups; ++g) {
        const size_t ng = gm[g].size();
        Dom* doms = (Dom*)(img.data() + off[g].doms);
        FinishDst* fin = (FinishDst*)(img.data() + off[g].fin);
        for (size_t i = 0; i < ng; ++i) {
            Dom role[ENS_ROLES];
            FinishDst f[3];
            member_views(members[gm[g][i]], role, f);
            for (int r = 0; r < ENS_ROLES; ++r) doms[r * ng + i] = role[r];
            for (int s3 = 0; s3 < 3; ++s3) fin[s3 * ng + i] = f[s3];
        }
        for (int k = 0; k < BK_COUNT; ++k) {
            int* start = (int*)(img.data() + off[g].start[k]);
            int* owner = (int*)(img.data() + off[g].owner[k]);
            int t = 0;
            for (size_t i = 0; i < ng; ++i) {
                start[i] = t;
                const int nb = batch_blocks(members[gm[g][i]], k, col_tile);
                for (int b = 0; b < nb; ++b) owner[t + b] = (int)i;
                t += nb;
            }
            start[ng] = t;
        }
    }
    ARW_API(ensemble)* e = new (std::nothrow) ARW_API(ensemble)();
    if (!e) return fail(ARW_E_ARG, "out of host memory");
    if (hipMalloc(&e->dev, bytes) != hipSuccess) {
        delete e;
        return fail(ARW_E_LAUNCH, "hipMalloc of the ensemble tables failed");
    }
    if (hipMemcpy(e->dev, img.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(e->dev);
        delete e;
        return fail(ARW_E_LAUNCH, "copy of the ensemble tables failed");
    }
    unsigned char* dev = (unsigned char*)e->dev;
    e->n = n;
    e->prm = *prm;
    e->b3 = members[0].base3d != 0;
    e->groups.resize(groups);
    for (int g = 0; g < groups; ++g) {
        const size_t ng = gm[g].size();
        for (int r = 0; r < ENS_ROLES; ++r) {
            const int stage = r < ENS_BASE ? r / 3 : -1;
            for (int k = 0; k < BK_COUNT; ++k) {
                Batch& b = e->groups[g].sets[r].k[k];
                b.doms = (const Dom*)(dev + off[g].doms) + r * ng;
                b.fin = (stage >= 0 && r % 3 == 0)
                            ? (const FinishDst*)(dev + off[g].fin) + stage * ng
                            : nullptr;
                b.owner = (const int*)(dev + off[g].owner[k]);
                b.start = (const int*)(dev + off[g].start[k]);
                b.total = off[g].total[k];
                b.tile = k == BK_COLT ? col_tile : 256;
            }
        }
    }
    *out = e;
    return ARW_OK;
}

int32_t ARW_API(ensemble_step)(ARW_API(ensemble)* e, int32_t nsteps,
                               void* stream)
{
    if (!e) return fail(ARW_E_ARG, "null ensemble");
    if (nsteps < 0) return fail(ARW_E_ARG, "nsteps must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    for (int t = 0; t < nsteps; ++t)
        for (const EnsGroup& G : e->groups) {
            int32_t rc = ens_issue(G, &e->prm, e->b3, st);
            if (rc) return rc;
        }
    return ARW_OK;
}

int32_t ARW_API(ensemble_capture)(ARW_API(ensemble)* e, int32_t nsteps,
                                  int32_t join_every)
{
    if (!e) return fail(ARW_E_ARG, "null ensemble");
    if (nsteps < 1) return fail(ARW_E_ARG, "nsteps must be >= 1");
    if (join_every < 0) return fail(ARW_E_ARG, "join_every must be >= 0");
    const int chunk = (join_every == 0 || join_every > nsteps) ? nsteps : join_every;
    if (e->exec) {
        hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    const int ng = (int)e->groups.size();
    std::vector<hipStream_t> cs(ng, nullptr);
    std::vector<hipEvent_t> ev(ng, nullptr);
    int32_t rc = ARW_OK;
    for (int g = 0; g < ng && rc == ARW_OK; ++g) {
        if (hipStreamCreateWithFlags(&cs[g], hipStreamNonBlocking) != hipSuccess
            || hipEventCreateWithFlags(&ev[g], hipEventDisableTiming) != hipSuccess)
            rc = fail(ARW_E_LAUNCH, "stream/event creation failed");
    }
    hipGraph_t graph = nullptr;
    bool capturing = false;
    if (rc == ARW_OK) {
        if (hipStreamBeginCapture(cs[0], hipStreamCaptureModeThreadLocal) != hipSuccess)
            rc = fail(ARW_E_LAUNCH, "hipStreamBeginCapture failed");
        else
            capturing = true;
    }
    for (int t0 = 0; t0 < nsteps && rc == ARW_OK; t0 += chunk) {
        const int nt = (nsteps - t0 < chunk) ? nsteps - t0 : chunk;
        // fork: every group's stream continues after cs[0]'s work so far
        if (ng > 1) {
            if (hipEventRecord(ev[0], cs[0]) != hipSuccess)
                rc = fail(ARW_E_LAUNCH, "fork record failed");
            for (int g = 1; g < ng && rc == ARW_OK; ++g)
                if (hipStreamWaitEvent(cs[g], ev[0], 0) != hipSuccess)
                    rc = fail(ARW_E_LAUNCH, "fork wait failed");
        }
        for (int g = 0; g < ng && rc == ARW_OK; ++g)
            for (int t = 0; t < nt && rc == ARW_OK; ++t)
                rc = ens_issue(e->groups[g], &e->prm, e->b3, cs[g]);
        // join
        for (int g = 1; g < ng && rc == ARW_OK; ++g)
            if (hipEventRecord(ev[g], cs[g]) != hipSuccess
                || hipStreamWaitEvent(cs[0], ev[g], 0) != hipSuccess)
                rc = fail(ARW_E_LAUNCH, "join failed");
    }
    if (capturing) {
        hipError_t er = hipStreamEndCapture(cs[0], &graph);
        if (rc == ARW_OK && er != hipSuccess) rc = fail(ARW_E_LAUNCH, "hipStreamEndCapture failed");
    }
    if (rc == ARW_OK && hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        e->exec = nullptr;
        rc = fail(ARW_E_LAUNCH, "hipGraphInstantiate failed");
    }
    if (graph) hipGraphDestroy(graph);
    for (int g = 0; g < ng; ++g) {
        if (ev[g]) hipEventDestroy(ev[g]);
        if (cs[g]) hipStreamDestroy(cs[g]);
    }
    return rc;
}

int32_t ARW_API(ensemble_launch)(ARW_API(ensemble)* e, void* stream)
{
    if (!e) return fail(ARW_E_ARG, "null ensemble");
    if (!e->exec) return fail(ARW_E_ARG, "no captured graph (call ensemble_capture)");
    if (hipGraphLaunch(e->exec, (hipStream_t)stream) != hipSuccess)
        return fail(ARW_E_LAUNCH, "hipGraphLaunch failed");
    return ARW_OK;
}

int32_t ARW_API(ensemble_destroy)(ARW_API(ensemble)* e)
{
    if (!e) return ARW_OK;
    if (e->exec) hipGraphExecDestroy(e->exec);
    if (e->dev) hipFree(e->dev);
    delete e;
    return ARW_OK;
}

int32_t ARW_API(profile_begin)(int32_t max_events)
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

int32_t ARW_API(profile_end)(double* ms, int32_t* launches, int32_t n_kernels)
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

const char* ARW_API(kernel_name)(int32_t id)
{
    return (id >= 0 && id < KID_COUNT) ? kKernelNames[id] : "";
}

int32_t ARW_API(kernel_count)(void) { return KID_COUNT; }

int32_t ARW_API(health)(const Dom* d, real* out, void* stream)
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

int32_t ARW_API(domain_mass)(const Dom* d, double* out, void* stream)
{
    int32_t rc = check_domain(d);
    if (rc) return rc;
    if (!out) return fail(ARW_E_ARG, "null output");
    hipLaunchKernelGGL(k_mass, dim3(1), dim3(1024), 0, (hipStream_t)stream,
                       *d, out);
    return launched("k_mass");
}

}  // extern "C"
