// WRF diff_opt = 2, km_opt = 4: the 2-D Smagorinsky held tendencies of a
// periodic domain with unit map factors (reference
// gpuwm/core/kernels/smag2d.cu:272-1340, driven once per step from the
// time-t fields by dycore.py:1223-1343 _compute_wrf_smag_tendencies).
//
// Passes, each a level-parallel point kernel in 64 x 4 workgroups:
//   deform    D11, D22 (mass), D12 (u-v corner), slope-corrected
//   km        K_m, K_h = K_m / Pr                 (needs D12 at 4 corners)
//   stress    tau11, tau22, tau12 in place of D (pointwise), tau13, tau23
//             (u / v faces of w levels), tau33 (mass)
//   u, v, w   stress divergences + vertical divergences + wall stress
//   fluxq     the scalar rows' face fluxes H1, H2: theta's t_2 (staged by
//             the metric pass) and each species, three rows per pass
//   divq      their divergences, theta's with the surface heat flux
// The stresses are stored once instead of being re-derived at every
// neighbour as the reference's on-demand device functions do; each
// stored value is the same expression in the same order, so the results
// are bit-identical.  The metrics (rdzw, rdz, zx, zy) are re-derived from
// the geopotential where they are read: a few loads and one division.
// Map factors are 1: each enters as a product or quotient with 1.0, which
// is exact, and is left out.
#include "march.h"

namespace ARW_NS {

namespace {

struct SmagSc {
    real rdx, rdy, dx, dy, cs, prandtl, cd0, heat;
    int vertical, drag, heatflux;
};

// scratch views; D11/D22/D12 become tau11/tau22/tau12 in place, and
// tau11/tau22's slots then hold the theta fluxes
struct SmagBuf {
    real *t11, *t22, *t12, *km, *kh, *t33, *t13, *t23;
    // the metrics of the time-t geopotential and the density, stored once
    real *rdzw, *rho, *rdz, *zx, *zy;
    real* t2;    // theta's t_2 = theta - 300
};

// periodic wrap of an index within [-1, n]: every neighbour these kernels
// read lies one cell (or the duplicate face) away
__device__ __forceinline__ int wrap(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

__device__ __forceinline__ int mix(const Dom& d, int k, int j, int i)
{
    return iM(k, wrap(j, d.ny), wrap(i, d.nx), d.ny, d.nx);
}

__device__ __forceinline__ real phi(const Dom& d, int kw, int j, int i)
{
    const int x = mix(d, kw, j, i);
    const real b = d.base3d ? d.phb[x] : d.phb[kw];
    return b + d.php[x];
}

// G / dz of mass level k, 2 G / (phi(k+1) - phi(k-1)) of w level kw
// (one-sided at the surface and the lid), the coordinate slopes at the
// u face i / v face j of w level kw (smag2d.cu:355-392)
__device__ __forceinline__ real rdzw_of(const Dom& d, int k, int j, int i)
{
    return G_F / (phi(d, k + 1, j, i) - phi(d, k, j, i));
}

__device__ __forceinline__ real rdz_of(const Dom& d, int kw, int j, int i)
{
    const real g2 = R(2.0) * G_F;
    if (kw <= 0) return g2 / (phi(d, 1, j, i) - phi(d, 0, j, i));
    if (kw >= d.nz) return g2 / (phi(d, d.nz, j, i) - phi(d, d.nz - 1, j, i));
    return g2 / (phi(d, kw + 1, j, i) - phi(d, kw - 1, j, i));
}

__device__ __forceinline__ real zx_of(const Dom& d, real rdx, int kw, int j,
                                      int i)
{
    return (rdx * (phi(d, kw, j, i) - phi(d, kw, j, i - 1))) / G_F;
}

__device__ __forceinline__ real zy_of(const Dom& d, real rdy, int kw, int j,
                                      int i)
{
    return (rdy * (phi(d, kw, j, i) - phi(d, kw, j - 1, i))) / G_F;
}

// density with the vapour loading, (1 + qv) / alt, and exactly 1 / alt
// for a dry state (smag2d.cu:394-404)
__device__ __forceinline__ real rho_of(const Dom& d, int k, int j, int i)
{
    const int x = mix(d, k, j, i);
    return d.qv ? (R(1.0) + d.qv[x]) / d.alt[x] : R(1.0) / d.alt[x];
}

// the stored metrics and density
__device__ __forceinline__ real rdzw(const Dom& d, const SmagBuf& b, int k,
                                     int j, int i)
{
    return b.rdzw[mix(d, k, j, i)];
}

__device__ __forceinline__ real rdz(const Dom& d, const SmagBuf& b, int kw,
                                    int j, int i)
{
    return b.rdz[mix(d, kw, j, i)];
}

__device__ __forceinline__ real zx(const Dom& d, const SmagBuf& b, int kw,
                                   int j, int i)
{
    return b.zx[mix(d, kw, j, i)];
}

__device__ __forceinline__ real zy(const Dom& d, const SmagBuf& b, int kw,
                                   int j, int i)
{
    return b.zy[mix(d, kw, j, i)];
}

__device__ __forceinline__ real rho(const Dom& d, const SmagBuf& b, int k,
                                    int j, int i)
{
    return b.rho[mix(d, k, j, i)];
}

__device__ __forceinline__ real U(const Dom& d, int k, int j, int i)
{
    return d.u[iU(k, wrap(j, d.ny), wrap(i, d.nx), d.ny, d.nx)];
}

__device__ __forceinline__ real V(const Dom& d, int k, int j, int i)
{
    return d.v[iV(k, wrap(j, d.ny), wrap(i, d.nx), d.ny, d.nx)];
}

__device__ __forceinline__ real W(const Dom& d, int k, int j, int i)
{
    return d.w[mix(d, k, j, i)];
}

__device__ __forceinline__ real at(const Dom& d, const real* a, int k, int j,
                                   int i)
{
    return a[mix(d, k, j, i)];
}

// 0.5 x WRF's w-level weights of the per-level sums p(k): the cf
// extrapolation at the surface, the one-sided cft2 form at the lid
// (smag2d.cu:420-431)
template <class P>
__device__ __forceinline__ real full(const Dom& d, int kw, P p)
{
    const int nz = d.nz;
    real r;
    if (kw <= 0) {
        r = (d.cf1 * p(0) + d.cf2 * p(1)) + d.cf3 * p(2);
    } else if (kw >= nz) {
        const real cft2 = (R(-0.5) * d.dnw[nz - 1]) / d.dn[nz - 1];
        r = (R(1.0) - cft2) * p(nz - 1) + cft2 * p(nz - 2);
    } else {
        r = d.fnm[kw] * p(kw) + d.fnp[kw] * p(kw - 1);
    }
    return R(0.5) * r;
}

// 0.5 (fnm (a_kw(s) + a_kw) + fnp (a_kw-1(s) + a_kw-1)) at interior w
// levels, 0 at the surface and the lid (smag2d.cu:688-733, :1206-1225);
// (sj, si) names the neighbour a(s)
__device__ __forceinline__ real mavg(const Dom& d, const real* a, int kw,
                                     int j, int i, int sj, int si)
{
    if (kw <= 0 || kw >= d.nz) return R(0.0);
    return R(0.5) * (d.fnm[kw] * (at(d, a, kw, j + sj, i + si)
                                  + at(d, a, kw, j, i))
                     + d.fnp[kw] * (at(d, a, kw - 1, j + sj, i + si)
                                    + at(d, a, kw - 1, j, i)));
}

// Each pass is a per-point body (k, j, i) launched one thread per point,
// 64 x 4 points of one level per workgroup.  (A column march of the same
// bodies, one thread per column looping over the levels, measured no
// faster and the theta flux pass 1.6x slower: without register carries
// the march only re-orders the same loads.)
#define SMAG_BODY(NAME, ...)                                                 \
    __device__ __forceinline__ void NAME(__VA_ARGS__, int k, int j, int i)
#define SMAG_POINTS(KERNEL, BODY, NXS, NYS, NLEV, PARAMS, ...)                \
    __global__ void KERNEL PARAMS                                            \
    {                                                                        \
        const int i = blockIdx.x * blockDim.x + threadIdx.x;                 \
        const int j = blockIdx.y * blockDim.y + threadIdx.y;                 \
        const int k = blockIdx.z;                                            \
        if (i >= (NXS) || j >= (NYS) || k >= (NLEV)) return;                \
        BODY(__VA_ARGS__, k, j, i);                                          \
    }

// rdzw and the density at mass level k, rdz, zx and zy at w level k
SMAG_BODY(smag_metric_at, const Dom& d, const SmagSc& c, const SmagBuf& b)
{
    const int x = iM(k, j, i, d.ny, d.nx);
    if (k < d.nz) {
        b.rdzw[x] = rdzw_of(d, k, j, i);
        b.rho[x] = rho_of(d, k, j, i);
        b.t2[x] = ((d.base3d ? d.thb[x] : d.thb[k]) + d.thp[x]) - R(300.0);
    }
    b.rdz[x] = rdz_of(d, k, j, i);
    b.zx[x] = zx_of(d, c.rdx, k, j, i);
    b.zy[x] = zy_of(d, c.rdy, k, j, i);
}
SMAG_POINTS(k_smag_metric, smag_metric_at, d.nx, d.ny, d.nz + 1,
           (Dom d, SmagSc c, SmagBuf b), d, c, b)

// WRF cal_deform_and_div (smag2d.cu:484-595)
SMAG_BODY(smag_deform_at, const Dom& d, const SmagSc& c, const SmagBuf& b)
{
    const real q = R(0.25);
    const real rdx = c.rdx, rdy = c.rdy;
    const int x = iM(k, j, i, d.ny, d.nx);
    const real rz = rdzw(d, b, k, j, i);
    {
        const real tmpzx = q * (((zx(d, b, k, j, i) + zx(d, b, k, j, i + 1))
                                 + zx(d, b, k + 1, j, i))
                                + zx(d, b, k + 1, j, i + 1));
        auto p = [&](int kk) { return U(d, kk, j, i) + U(d, kk, j, i + 1); };
        const real slope = ((full(d, k + 1, p) - full(d, k, p)) * tmpzx) * rz;
        b.t11[x] = R(2.0) * (rdx * (U(d, k, j, i + 1) - U(d, k, j, i)) - slope);
    }
    {
        const real tmpzy = q * (((zy(d, b, k, j, i) + zy(d, b, k, j + 1, i))
                                 + zy(d, b, k + 1, j, i))
                                + zy(d, b, k + 1, j + 1, i));
        auto p = [&](int kk) { return V(d, kk, j, i) + V(d, kk, j + 1, i); };
        const real slope = ((full(d, k + 1, p) - full(d, k, p)) * tmpzy) * rz;
        b.t22[x] = R(2.0) * (rdy * (V(d, k, j + 1, i) - V(d, k, j, i)) - slope);
    }
    const real rr = ((rz + rdzw(d, b, k, j, i - 1)) + rdzw(d, b, k, j - 1, i - 1))
                    + rdzw(d, b, k, j - 1, i);
    const real tzy = q * (((zy(d, b, k, j, i - 1) + zy(d, b, k, j, i))
                           + zy(d, b, k + 1, j, i - 1))
                          + zy(d, b, k + 1, j, i));
    auto pu = [&](int kk) { return U(d, kk, j - 1, i) + U(d, kk, j, i); };
    const real uslope = (((full(d, k + 1, pu) - full(d, k, pu)) * q) * tzy) * rr;
    const real tzx = q * (((zx(d, b, k, j - 1, i) + zx(d, b, k, j, i))
                           + zx(d, b, k + 1, j - 1, i))
                          + zx(d, b, k + 1, j, i));
    auto pv = [&](int kk) { return V(d, kk, j, i - 1) + V(d, kk, j, i); };
    const real vslope = (((full(d, k + 1, pv) - full(d, k, pv)) * q) * tzx) * rr;
    b.t12[x] = ((rdy * (U(d, k, j, i) - U(d, k, j - 1, i)) - uslope)
                + rdx * (V(d, k, j, i) - V(d, k, j, i - 1))) - vslope;
}
SMAG_POINTS(k_smag_deform, smag_deform_at, d.nx, d.ny, d.nz,
           (Dom d, SmagSc c, SmagBuf b), d, c, b)

// WRF smag2d_km (smag2d.cu:604-652)
SMAG_BODY(smag_km_at, const Dom& d, const SmagSc& c, const SmagBuf& b)
{
    const real q = R(0.25);
    const int x = iM(k, j, i, d.ny, d.nx);
    const real d11 = b.t11[x], d22 = b.t22[x];
    const real d12 = q * (((at(d, b.t12, k, j, i) + at(d, b.t12, k, j + 1, i))
                           + at(d, b.t12, k, j, i + 1))
                          + at(d, b.t12, k, j + 1, i + 1));
    const real strain = sqrt((q * (d11 - d22)) * (d11 - d22) + d12 * d12);
    const real mlen = sqrt(c.dx * c.dy);
    real km = fmin((((c.cs * c.cs) * mlen) * mlen) * strain, R(10.0) * mlen);
    const real rz = rdzw(d, b, k, j, i);
    const real sx = ((q * (((fabs(zx(d, b, k, j, i))
                             + fabs(zx(d, b, k, j, i + 1)))
                            + fabs(zx(d, b, k + 1, j, i)))
                           + fabs(zx(d, b, k + 1, j, i + 1)))) * rz) * c.dx;
    const real sy = ((q * (((fabs(zy(d, b, k, j, i))
                             + fabs(zy(d, b, k, j + 1, i)))
                            + fabs(zy(d, b, k + 1, j, i)))
                           + fabs(zy(d, b, k + 1, j + 1, i)))) * rz) * c.dy;
    const real alpha = fmax(sqrt(sx * sx + sy * sy), R(1.0));
    const real limit = fmax(R(10.0) / mlen, R(1.0e-3));
    km = km / (strain > limit ? alpha * alpha : alpha);
    b.km[x] = km;
    b.kh[x] = km / c.prandtl;
}
SMAG_POINTS(k_smag_km, smag_km_at, d.nx, d.ny, d.nz,
           (Dom d, SmagSc c, SmagBuf b), d, c, b)

// tau11, tau22, tau12 = -rho K D in place (smag2d.cu:659-686)
SMAG_BODY(smag_tau_at, const Dom& d, const SmagBuf& b)
{
    const real q = R(0.25);
    const int x = iM(k, j, i, d.ny, d.nx);
    const real r = rho(d, b, k, j, i), km = b.km[x];
    b.t11[x] = ((-r) * km) * b.t11[x];
    b.t22[x] = ((-r) * km) * b.t22[x];
    const real ra = q * (((r + rho(d, b, k, j, i - 1)) + rho(d, b, k, j - 1, i - 1))
                         + rho(d, b, k, j - 1, i));
    const real ka = q * (((km + at(d, b.km, k, j, i - 1))
                          + at(d, b.km, k, j - 1, i - 1))
                         + at(d, b.km, k, j - 1, i));
    b.t12[x] = ((-ra) * ka) * b.t12[x];
}

// tau13 / tau23 at the u / v faces of w level kw (0 at the surface and the
// lid) and tau33 at mass level kw (smag2d.cu:826-899, :945-956)
SMAG_BODY(smag_tauw_at, const Dom& d, const SmagSc& c, const SmagBuf& b)
{
    const int kw = k, nz = d.nz;
    const real q = R(0.25), h = R(0.5);
    const int x = iM(kw, j, i, d.ny, d.nx);
    if (kw < nz) {
        const real dw = (R(2.0) * (W(d, kw + 1, j, i) - W(d, kw, j, i)))
                        * rdzw(d, b, kw, j, i);
        b.t33[x] = ((-rho(d, b, kw, j, i)) * b.km[x]) * dw;
    }
    if (kw == 0 || kw == nz) {
        b.t13[x] = R(0.0);
        b.t23[x] = R(0.0);
        return;
    }
    const real fm = d.fnm[kw], fp = d.fnp[kw];
    {
        const real rz = h * (rdz(d, b, kw, j, i) + rdz(d, b, kw, j, i - 1));
        auto wa = [&](int kk) {
            return q * (((W(d, kk, j, i) + W(d, kk + 1, j, i))
                         + W(d, kk, j, i - 1)) + W(d, kk + 1, j, i - 1));
        };
        const real slope = ((wa(kw) - wa(kw - 1)) * zx(d, b, kw, j, i)) * rz;
        const real dw = c.rdx * (W(d, kw, j, i) - W(d, kw, j, i - 1)) - slope;
        const real d13 = dw + (U(d, kw, j, i) - U(d, kw - 1, j, i)) * rz;
        const real ra = h * (fm * (rho(d, b, kw, j, i - 1) + rho(d, b, kw, j, i))
                             + fp * (rho(d, b, kw - 1, j, i - 1)
                                     + rho(d, b, kw - 1, j, i)));
        const real ka = h * (fm * (at(d, b.km, kw, j, i - 1)
                                   + at(d, b.km, kw, j, i))
                             + fp * (at(d, b.km, kw - 1, j, i - 1)
                                     + at(d, b.km, kw - 1, j, i)));
        b.t13[x] = ((-ra) * ka) * d13;
    }
    {
        const real rz = h * (rdz(d, b, kw, j, i) + rdz(d, b, kw, j - 1, i));
        auto wa = [&](int kk) {
            return q * (((W(d, kk, j, i) + W(d, kk + 1, j, i))
                         + W(d, kk, j - 1, i)) + W(d, kk + 1, j - 1, i));
        };
        const real slope = ((wa(kw) - wa(kw - 1)) * zy(d, b, kw, j, i)) * rz;
        const real dw = c.rdy * (W(d, kw, j, i) - W(d, kw, j - 1, i)) - slope;
        const real d23 = dw + (V(d, kw, j, i) - V(d, kw - 1, j, i)) * rz;
        const real ra = h * (fm * (rho(d, b, kw, j - 1, i) + rho(d, b, kw, j, i))
                             + fp * (rho(d, b, kw - 1, j - 1, i)
                                     + rho(d, b, kw - 1, j, i)));
        const real ka = h * (fm * (at(d, b.km, kw, j - 1, i)
                                   + at(d, b.km, kw, j, i))
                             + fp * (at(d, b.km, kw - 1, j - 1, i)
                                     + at(d, b.km, kw - 1, j, i)));
        b.t23[x] = ((-ra) * ka) * d23;
    }
}
// one launch for every stress: the horizontal ones of mass level k (in
// place of the deformation) and the vertical ones of w level k
SMAG_BODY(smag_stress_at, const Dom& d, const SmagSc& c, const SmagBuf& b)
{
    if (k < d.nz) smag_tau_at(d, b, k, j, i);
    smag_tauw_at(d, c, b, k, j, i);
}
SMAG_POINTS(k_smag_stress, smag_stress_at, d.nx, d.ny, d.nz + 1,
           (Dom d, SmagSc c, SmagBuf b), d, c, b)

// u: horizontal_diffusion_u_2 (smag2d.cu:765-795), vertical_diffusion_u_2
// (:916-928) and the isfflx = 0 wall stress (:1608-1628); every u face,
// the periodic duplicate included
SMAG_BODY(smag_u_at, const Dom& d, const SmagSc& c, const SmagBuf& b, real* __restrict__ out)
{
    const real q = R(0.25), h = R(0.5), e = R(0.125), one = R(1.0);
    const real rdx = c.rdx, rdy = c.rdy;
    const real tmpdz = h * (one / rdzw(d, b, k, j, i) + one / rdzw(d, b, k, j, i - 1));
    const real zxu = h * (zx(d, b, k, j, i) + zx(d, b, k + 1, j, i));
    const real zyu = e * (((((((zy(d, b, k, j, i - 1) + zy(d, b, k, j, i))
                               + zy(d, b, k, j + 1, i - 1))
                              + zy(d, b, k, j + 1, i))
                             + zy(d, b, k + 1, j, i - 1))
                            + zy(d, b, k + 1, j, i))
                           + zy(d, b, k + 1, j + 1, i - 1))
                          + zy(d, b, k + 1, j + 1, i));
    const real divh = rdx * (at(d, b.t11, k, j, i) - at(d, b.t11, k, j, i - 1))
                      + rdy * (at(d, b.t12, k, j + 1, i) - at(d, b.t12, k, j, i));
    const real divz = (zxu * (mavg(d, b.t11, k + 1, j, i, 0, -1)
                              - mavg(d, b.t11, k, j, i, 0, -1))) / tmpdz
                      + (zyu * (mavg(d, b.t12, k + 1, j, i, 1, 0)
                                - mavg(d, b.t12, k, j, i, 1, 0))) / tmpdz;
    real r = R(0.0) + ((G_F * tmpdz) / d.dnw[k]) * (divh - divz);
    if (c.vertical) {
        r = r + (G_F * (at(d, b.t13, k + 1, j, i) - at(d, b.t13, k, j, i)))
                / d.dnw[k];
        if (c.drag && k == 0) {
            const real vv = q * (((V(d, 0, j, i) + V(d, 0, j + 1, i))
                                  + V(d, 0, j, i - 1)) + V(d, 0, j + 1, i - 1));
            const real uu = U(d, 0, j, i);
            const real speed = sqrt(uu * uu + vv * vv) + R(1.0e-15);
            const real stress = (c.cd0 * speed) * uu;
            const real ra = h * (rho(d, b, 0, j, i) + rho(d, b, 0, j, i - 1));
            r = r + ((G_F * stress) * ra) / d.dnw[0];
        }
    }
    out[iU(k, j, i, d.ny, d.nx)] = r;
}
SMAG_POINTS(k_smag_u, smag_u_at, d.nx + 1, d.ny, d.nz,
           (Dom d, SmagSc c, SmagBuf b, real* __restrict__ out), d, c, b, out)

// v: horizontal_diffusion_v_2 (smag2d.cu:798-824), vertical_diffusion_v_2
// (:931-943) and the isfflx = 0 wall stress (:1630-1654)
SMAG_BODY(smag_v_at, const Dom& d, const SmagSc& c, const SmagBuf& b, real* __restrict__ out)
{
    const real q = R(0.25), h = R(0.5), e = R(0.125), one = R(1.0);
    const real rdx = c.rdx, rdy = c.rdy;
    const real tmpdz = h * (one / rdzw(d, b, k, j, i) + one / rdzw(d, b, k, j - 1, i));
    const real zxv = e * (((((((zx(d, b, k, j - 1, i) + zx(d, b, k, j - 1, i + 1))
                               + zx(d, b, k, j, i))
                              + zx(d, b, k, j, i + 1))
                             + zx(d, b, k + 1, j - 1, i))
                            + zx(d, b, k + 1, j - 1, i + 1))
                           + zx(d, b, k + 1, j, i))
                          + zx(d, b, k + 1, j, i + 1));
    const real zyv = h * (zy(d, b, k, j, i) + zy(d, b, k + 1, j, i));
    const real divh = rdy * (at(d, b.t22, k, j, i) - at(d, b.t22, k, j - 1, i))
                      + rdx * (at(d, b.t12, k, j, i + 1) - at(d, b.t12, k, j, i));
    const real divz = (zxv * (mavg(d, b.t12, k + 1, j, i, 0, 1)
                              - mavg(d, b.t12, k, j, i, 0, 1))) / tmpdz
                      + (zyv * (mavg(d, b.t22, k + 1, j, i, -1, 0)
                                - mavg(d, b.t22, k, j, i, -1, 0))) / tmpdz;
    real r = R(0.0) + ((G_F * tmpdz) / d.dnw[k]) * (divh - divz);
    if (c.vertical) {
        r = r + (G_F * (at(d, b.t23, k + 1, j, i) - at(d, b.t23, k, j, i)))
                / d.dnw[k];
        if (c.drag && k == 0) {
            const real uu = q * (((U(d, 0, j, i) + U(d, 0, j, i + 1))
                                  + U(d, 0, j - 1, i)) + U(d, 0, j - 1, i + 1));
            const real vv = V(d, 0, j, i);
            const real speed = sqrt(vv * vv + uu * uu) + R(1.0e-15);
            const real stress = (c.cd0 * speed) * vv;
            const real ra = h * (rho(d, b, 0, j, i) + rho(d, b, 0, j - 1, i));
            r = r + ((G_F * stress) * ra) / d.dnw[0];
        }
    }
    out[iV(k, j, i, d.ny, d.nx)] = r;
}
SMAG_POINTS(k_smag_v, smag_v_at, d.nx, d.ny + 1, d.nz,
           (Dom d, SmagSc c, SmagBuf b, real* __restrict__ out), d, c, b, out)

// w: horizontal_diffusion_w_2 (smag2d.cu:1071-1093) and
// vertical_diffusion_w_2 (:958-973); 0 at the surface and the lid
SMAG_BODY(smag_w_at, const Dom& d, const SmagSc& c, const SmagBuf& b, real* __restrict__ out)
{
    const int kw = k, nz = d.nz;
    const int x = iM(kw, j, i, d.ny, d.nx);
    if (kw == 0 || kw == nz) {
        out[x] = R(0.0);
        return;
    }
    const real q = R(0.25), h = R(0.5);
    const real zxw = h * (zx(d, b, kw, j, i) + zx(d, b, kw, j, i + 1));
    const real zyw = h * (zy(d, b, kw, j, i) + zy(d, b, kw, j + 1, i));
    const real rz = rdz(d, b, kw, j, i);
    const real divh = c.rdx * (at(d, b.t13, kw, j, i + 1) - at(d, b.t13, kw, j, i))
                      + c.rdy * (at(d, b.t23, kw, j + 1, i) - at(d, b.t23, kw, j, i));
    auto m13 = [&](int kk) {
        return q * (((at(d, b.t13, kk, j, i) + at(d, b.t13, kk, j, i + 1))
                     + at(d, b.t13, kk + 1, j, i)) + at(d, b.t13, kk + 1, j, i + 1));
    };
    auto m23 = [&](int kk) {
        return q * (((at(d, b.t23, kk, j, i) + at(d, b.t23, kk, j + 1, i))
                     + at(d, b.t23, kk + 1, j, i)) + at(d, b.t23, kk + 1, j + 1, i));
    };
    const real divz = rz * (zxw * (m13(kw) - m13(kw - 1))
                            + zyw * (m23(kw) - m23(kw - 1)));
    real r = R(0.0) + (G_F / (d.dn[kw] * rz)) * (divh - divz);
    if (c.vertical)
        r = r + (G_F * (b.t33[x] - at(d, b.t33, kw - 1, j, i))) / d.dn[kw];
    out[x] = r;
}
SMAG_POINTS(k_smag_w, smag_w_at, d.nx, d.ny, d.nz + 1,
           (Dom d, SmagSc c, SmagBuf b, real* __restrict__ out), d, c, b, out)

// The scalar rows -- theta's t_2 = theta - 300 (staged by the metric pass
// from the stored perturbation, smag2d.cu:1105-1119) and every species'
// mixing ratio -- in groups of up to three, the work area's free slots
// holding their fluxes: one pass forms each row's face fluxes H1 (u face
// i), H2 (v face j) (smag2d.cu:1164-1200, wrf_smag_flux_s :1236-1251),
// sharing the density, K_h and metric loads of the point, and one pass
// their divergences (wrf_smag_hd_s :1254-1292), theta's with the constant
// surface heat flux under isfflx 0 / 2 (:1658-1668).
struct SmagQ {
    const real* src[3];
    real* fx[3];
    real* fy[3];
    real* out[3];
    int heat0;               // row 0 is theta: add the surface heat flux
};

template <int NS>
__global__ void k_smag_fluxq(Dom d, SmagSc c, SmagBuf b, SmagQ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny || k >= d.nz) return;
    const real h = R(0.5), one = R(1.0), two = R(2.0);
    const int x = iM(k, j, i, d.ny, d.nx);
    const real rz = rdzw(d, b, k, j, i);
    const real rax = h * (rho(d, b, k, j, i - 1) + rho(d, b, k, j, i));
    const real kax = h * (at(d, b.kh, k, j, i - 1) + b.kh[x]);
    const real tzx = h * (zx(d, b, k, j, i) + zx(d, b, k + 1, j, i));
    const real rdzu = two / (one / rz + one / rdzw(d, b, k, j, i - 1));
    const real ray = h * (rho(d, b, k, j - 1, i) + rho(d, b, k, j, i));
    const real kay = h * (at(d, b.kh, k, j - 1, i) + b.kh[x]);
    const real tzy = h * (zy(d, b, k, j, i) + zy(d, b, k + 1, j, i));
    const real rdzv = two / (one / rz + one / rdzw(d, b, k, j - 1, i));
#pragma unroll
    for (int e = 0; e < NS; ++e) {
        const real* S = q.src[e];
        auto sv = [&](int kk, int jj, int ii) { return S[mix(d, kk, jj, ii)]; };
        const real s0 = S[x];
        auto px = [&](int kk) { return sv(kk, j, i - 1) + sv(kk, j, i); };
        const real gx = c.rdx * (s0 - sv(k, j, i - 1))
                        - (tzx * (full(d, k + 1, px) - full(d, k, px))) * rdzu;
        q.fx[e][x] = ((-rax) * kax) * gx;
        auto py = [&](int kk) { return sv(kk, j - 1, i) + sv(kk, j, i); };
        const real gy = c.rdy * (s0 - sv(k, j - 1, i))
                        - (tzy * (full(d, k + 1, py) - full(d, k, py))) * rdzv;
        q.fy[e][x] = ((-ray) * kay) * gy;
    }
}

template <int NS>
__global__ void k_smag_divq(Dom d, SmagSc c, SmagBuf b, SmagQ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= d.nx || j >= d.ny || k >= d.nz) return;
    const real qq = R(0.25);
    const int x = iM(k, j, i, d.ny, d.nx);
    const real zxm = qq * (((zx(d, b, k, j, i) + zx(d, b, k, j, i + 1))
                            + zx(d, b, k + 1, j, i)) + zx(d, b, k + 1, j, i + 1));
    const real zym = qq * (((zy(d, b, k, j, i) + zy(d, b, k, j + 1, i))
                            + zy(d, b, k + 1, j, i)) + zy(d, b, k + 1, j + 1, i));
    const real rz = rdzw(d, b, k, j, i);
    const real fac = G_F / (d.dnw[k] * rz);
#pragma unroll
    for (int e = 0; e < NS; ++e) {
        const real* fx = q.fx[e];
        const real* fy = q.fy[e];
        const real divh = c.rdx * (at(d, fx, k, j, i + 1) - fx[x])
                          + c.rdy * (at(d, fy, k, j + 1, i) - fy[x]);
        const real divz = (zxm * (mavg(d, fx, k + 1, j, i, 0, 1)
                                  - mavg(d, fx, k, j, i, 0, 1))) * rz
                          + (zym * (mavg(d, fy, k + 1, j, i, 1, 0)
                                    - mavg(d, fy, k, j, i, 1, 0))) * rz;
        real r = R(0.0) + fac * (divh - divz);
        if (e == 0 && q.heat0 && c.heatflux && k == 0)
            r = r - ((G_F * c.heat) * rho(d, b, 0, j, i)) / d.dnw[0];
        q.out[e][x] = r;
    }
}

}  // namespace

// The held km_opt = 4 tendencies of the time-t fields of d (u, v, w, thp,
// php, the EOS's alt and, moist, qv) into out_u, out_v, out_w, out_t and
// of the nq species q_src[e] into held_q + e nz ny nx, written (not
// added); buf holds 9 nz ny nx + 5 (nz+1) ny nx elements of scratch.
hipError_t run_smag(const Dom& d, const arw_params* prm, real* buf,
                    real* out_u, real* out_v, real* out_w, real* out_t,
                    const real* const* q_src, real* held_q, int nq,
                    hipStream_t st)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const size_t n = (size_t)nz * ny * nx, nw = (size_t)(nz + 1) * ny * nx;
    SmagBuf b;
    b.t11 = buf;
    b.t22 = buf + n;
    b.t12 = buf + 2 * n;
    b.km = buf + 3 * n;
    b.kh = buf + 4 * n;
    b.t33 = buf + 5 * n;
    b.rdzw = buf + 6 * n;
    b.rho = buf + 7 * n;
    b.t13 = buf + 8 * n;
    b.t23 = buf + 8 * n + nw;
    b.rdz = buf + 8 * n + 2 * nw;
    b.zx = buf + 8 * n + 3 * nw;
    b.zy = buf + 8 * n + 4 * nw;
    b.t2 = buf + 8 * n + 5 * nw;
    SmagSc c;
    c.rdx = (real)(1.0 / prm->dx);
    c.rdy = (real)(1.0 / prm->dy);
    c.dx = (real)prm->dx;
    c.dy = (real)prm->dy;
    c.cs = (real)prm->c_s;
    c.prandtl = (real)(1.0 / 3.0);
    c.cd0 = (real)prm->tke_drag_coefficient;
    c.heat = (real)prm->tke_heat_flux;
    c.vertical = prm->bl_pbl_physics == 0;
    c.drag = c.vertical && prm->isfflx == 0;
    c.heatflux = c.vertical && (prm->isfflx == 0 || prm->isfflx == 2);
    const dim3 blk(64, 4);
    auto grid = [&](int nxs, int nys, int nlev) {
        return dim3((nxs + 63) / 64, (nys + 3) / 4, nlev);
    };
    hipLaunchKernelGGL(k_smag_metric, grid(nx, ny, nz + 1), blk, 0, st, d, c, b);
    hipLaunchKernelGGL(k_smag_deform, grid(nx, ny, nz), blk, 0, st, d, c, b);
    hipLaunchKernelGGL(k_smag_km, grid(nx, ny, nz), blk, 0, st, d, c, b);
    hipLaunchKernelGGL(k_smag_stress, grid(nx, ny, nz + 1), blk, 0, st, d, c, b);
    hipLaunchKernelGGL(k_smag_u, grid(nx + 1, ny, nz), blk, 0, st, d, c, b, out_u);
    hipLaunchKernelGGL(k_smag_v, grid(nx, ny + 1, nz), blk, 0, st, d, c, b, out_v);
    hipLaunchKernelGGL(k_smag_w, grid(nx, ny, nz + 1), blk, 0, st, d, c, b, out_w);
    // theta and the species in groups of three; their fluxes take the
    // work area's slots the momentum passes are done with
    real* fxs[3] = {b.t11, b.t12, b.t13};
    real* fys[3] = {b.t22, b.t33, b.t23};
    const int nr = 1 + nq;                      // row 0: theta
    auto src_of = [&](int r) { return r == 0 ? (const real*)b.t2 : q_src[r - 1]; };
    auto out_of = [&](int r) { return r == 0 ? out_t : held_q + (size_t)(r - 1) * n; };
    for (int e0 = 0; e0 < nr; e0 += 3) {
        SmagQ q;
        const int ns = nr - e0 < 3 ? nr - e0 : 3;
        for (int e = 0; e < 3; ++e) {
            const int r = e0 + (e < ns ? e : 0);
            q.src[e] = src_of(r);
            q.fx[e] = fxs[e];
            q.fy[e] = fys[e];
            q.out[e] = out_of(r);
        }
        q.heat0 = e0 == 0;
        const dim3 g = grid(nx, ny, nz);
        if (ns == 3) {
            hipLaunchKernelGGL(k_smag_fluxq<3>, g, blk, 0, st, d, c, b, q);
            hipLaunchKernelGGL(k_smag_divq<3>, g, blk, 0, st, d, c, b, q);
        } else if (ns == 2) {
            hipLaunchKernelGGL(k_smag_fluxq<2>, g, blk, 0, st, d, c, b, q);
            hipLaunchKernelGGL(k_smag_divq<2>, g, blk, 0, st, d, c, b, q);
        } else {
            hipLaunchKernelGGL(k_smag_fluxq<1>, g, blk, 0, st, d, c, b, q);
            hipLaunchKernelGGL(k_smag_divq<1>, g, blk, 0, st, d, c, b, q);
        }
    }
    return hipGetLastError();
}

}  // namespace ARW_NS
