// Hand-written adjoints of the flux-form advection operators (periodic,
// msf = 1): flux_div_scalar / _u / _v / _w of the reference
// (advection.cu:206-416, 424-611, 616-794, 803-1051).
//
// Every operator sums face fluxes F_f = sum_o w_o(vel_f) q[s_f + o], with
// w_o(v) = C_o v + S_o |v| (flux5 horizontally, the Omega-signed flux3
// vertically) or a 2nd-order average next to the vertical boundaries, and
// vel_f a fixed average of the mass fluxes.  For an output adjoint lam the
// transpose is, with Lambda_f the adjoint of face flux f (the scaled
// difference of the output adjoints of the two cells it separates):
//   dL/dq[c]   = sum over faces f reading c of w_o(vel_f) Lambda_f
//   dL/dvel_f  = Lambda_f sum_o (C_o + S_o sign(vel_f)) q[s_f + o]
//   dL/dru ... = the transpose of the velocity averages applied to dL/dvel.
// sign(0) = 0 (a subgradient of |v|).  One thread per point of the union of
// the staggers gathers all four gradients at its indices, so there are no
// atomics and every output is written by exactly one thread.
//
// Faces along a horizontal line of n cells (face f between cells f-1 and
// f): "dup" lines have faces 0..n, face n being a separate copy of face 0
// with its own velocity whose flux enters only cell n-1 (the scalar, v-x,
// u-y and w lines, as in the forward, which reads the duplicate column /
// row of ru / rv); "periodic" lines have n faces (u-x, v-y: faces at mass
// centres, velocity averaged from two mass fluxes).
#include "common.h"
#include "march.h"

namespace ARW_NS {

namespace {

struct Args {
    const real *q, *ru, *rv, *rw, *lam, *rdz, *fnm, *fnp;
    real *gq, *gru, *grv, *grw;
    real rdx, rdy;
    int nz, ny, nx;
};

__device__ __forceinline__ real sgn(real v)
{
    return v > R(0.0) ? R(1.0) : (v < R(0.0) ? R(-1.0) : R(0.0));
}

// flux5 weights of qm3..qp2 and flux3 weights of qm2..qp1:
// w_o(v) = (C_o v + S_o |v|) / den.
__device__ __forceinline__ real c5(int o)
{
    constexpr real c[6] = {R(1.0), R(-8.0), R(37.0), R(37.0), R(-8.0),
                           R(1.0)};
    return c[o] / R(60.0);
}
__device__ __forceinline__ real s5(int o)
{
    constexpr real s[6] = {R(1.0), R(-5.0), R(10.0), R(-10.0), R(5.0),
                           R(-1.0)};
    return s[o] / R(60.0);
}
__device__ __forceinline__ real c3(int o)
{
    constexpr real c[4] = {R(-1.0), R(7.0), R(7.0), R(-1.0)};
    return c[o] / R(12.0);
}
__device__ __forceinline__ real s3(int o)
{
    constexpr real s[4] = {R(1.0), R(-3.0), R(3.0), R(-1.0)};
    return s[o] / R(12.0);
}

// ---------------------------------------------------------------------------
// Operator descriptions: the output adjoint folded onto the field's cells
// (lam), the field (q), and the face velocities of each direction.
// ---------------------------------------------------------------------------

// scalar at mass points (reference advection.cu:384-415)
struct OpScalar {
    static constexpr bool dupx = true, dupy = true, wlev = false;
    static __device__ real lam(const Args& a, int k, int j, int i)
    {
        return a.lam[iM(k, j, i, a.ny, a.nx)];
    }
    static __device__ real q(const Args& a, int k, int j, int i)
    {
        return a.q[iM(k, j, i, a.ny, a.nx)];
    }
    static __device__ int qidx(const Args& a, int k, int j, int i)
    {
        return iM(k, j, i, a.ny, a.nx);
    }
    static __device__ real velx(const Args& a, int k, int j, int f)
    {
        return a.ru[iU(k, j, f, a.ny, a.nx)];
    }
    static __device__ real vely(const Args& a, int k, int g, int i)
    {
        return a.rv[iV(k, g, i, a.ny, a.nx)];
    }
    static __device__ real velz(const Args& a, int kf, int j, int i)
    {
        return a.rw[iM(kf, j, i, a.ny, a.nx)];
    }
};

// u at u points, column nx duplicating column 0 (advection.cu:572-611)
struct OpU {
    static constexpr bool dupx = false, dupy = true, wlev = false;
    static __device__ real lam(const Args& a, int k, int j, int i)
    {
        real l = a.lam[iU(k, j, i, a.ny, a.nx)];
        if (i == 0) l += a.lam[iU(k, j, a.nx, a.ny, a.nx)];
        return l;
    }
    static __device__ real q(const Args& a, int k, int j, int i)
    {
        return a.q[iU(k, j, i, a.ny, a.nx)];
    }
    static __device__ int qidx(const Args& a, int k, int j, int i)
    {
        return iU(k, j, i, a.ny, a.nx);
    }
    // periodic face f sits at mass centre m = f - 1 (between u[m], u[m+1])
    static __device__ real velx(const Args& a, int k, int j, int f)
    {
        const int m = (f == 0) ? a.nx - 1 : f - 1;
        return R(0.5) * (a.ru[iU(k, j, m, a.ny, a.nx)]
                         + a.ru[iU(k, j, m + 1, a.ny, a.nx)]);
    }
    static __device__ real vely(const Args& a, int k, int g, int i)
    {
        const int im = (i == 0) ? a.nx - 1 : i - 1;
        return R(0.5) * (a.rv[iV(k, g, im, a.ny, a.nx)]
                         + a.rv[iV(k, g, i, a.ny, a.nx)]);
    }
    static __device__ real velz(const Args& a, int kf, int j, int i)
    {
        const int im = (i == 0) ? a.nx - 1 : i - 1;
        return R(0.5) * (a.rw[iM(kf, j, im, a.ny, a.nx)]
                         + a.rw[iM(kf, j, i, a.ny, a.nx)]);
    }
};

// v at v points, row ny duplicating row 0 (advection.cu:754-794)
struct OpV {
    static constexpr bool dupx = true, dupy = false, wlev = false;
    static __device__ real lam(const Args& a, int k, int j, int i)
    {
        real l = a.lam[iV(k, j, i, a.ny, a.nx)];
        if (j == 0) l += a.lam[iV(k, a.ny, i, a.ny, a.nx)];
        return l;
    }
    static __device__ real q(const Args& a, int k, int j, int i)
    {
        return a.q[iV(k, j, i, a.ny, a.nx)];
    }
    static __device__ int qidx(const Args& a, int k, int j, int i)
    {
        return iV(k, j, i, a.ny, a.nx);
    }
    static __device__ real velx(const Args& a, int k, int j, int f)
    {
        const int jm = (j == 0) ? a.ny - 1 : j - 1;
        return R(0.5) * (a.ru[iU(k, jm, f, a.ny, a.nx)]
                         + a.ru[iU(k, j, f, a.ny, a.nx)]);
    }
    static __device__ real vely(const Args& a, int k, int g, int i)
    {
        const int m = (g == 0) ? a.ny - 1 : g - 1;
        return R(0.5) * (a.rv[iV(k, m, i, a.ny, a.nx)]
                         + a.rv[iV(k, m + 1, i, a.ny, a.nx)]);
    }
    static __device__ real velz(const Args& a, int kf, int j, int i)
    {
        const int jm = (j == 0) ? a.ny - 1 : j - 1;
        return R(0.5) * (a.rw[iM(kf, jm, i, a.ny, a.nx)]
                         + a.rw[iM(kf, j, i, a.ny, a.nx)]);
    }
};

// w at w levels; the operator's rows 0 and nz are zero
// (advection.cu:1016-1050)
struct OpW {
    static constexpr bool dupx = true, dupy = true, wlev = true;
    static __device__ real lam(const Args& a, int k, int j, int i)
    {
        if (k == 0 || k == a.nz) return R(0.0);
        return a.lam[iM(k, j, i, a.ny, a.nx)];
    }
    static __device__ real q(const Args& a, int k, int j, int i)
    {
        return a.q[iM(k, j, i, a.ny, a.nx)];
    }
    static __device__ int qidx(const Args& a, int k, int j, int i)
    {
        return iM(k, j, i, a.ny, a.nx);
    }
    static __device__ real velx(const Args& a, int k, int j, int f)
    {
        return a.fnm[k] * a.ru[iU(k, j, f, a.ny, a.nx)]
               + a.fnp[k] * a.ru[iU(k - 1, j, f, a.ny, a.nx)];
    }
    static __device__ real vely(const Args& a, int k, int g, int i)
    {
        return a.fnm[k] * a.rv[iV(k, g, i, a.ny, a.nx)]
               + a.fnp[k] * a.rv[iV(k - 1, g, i, a.ny, a.nx)];
    }
    // face at mass level m, between w[m] and w[m+1]
    static __device__ real velz(const Args& a, int m, int j, int i)
    {
        return R(0.5) * (a.rw[iM(m, j, i, a.ny, a.nx)]
                         + a.rw[iM(m + 1, j, i, a.ny, a.nx)]);
    }
};

// ---------------------------------------------------------------------------
// Horizontal directions.  The x line of (k, j) and the y line of (k, i)
// are written once each; DUP selects the face layout.
// ---------------------------------------------------------------------------

template <class OP>
__device__ real lam_x(const Args& a, int k, int j, int f)
{
    const int n = a.nx;
    if (OP::dupx) {
        real r = (f < n) ? OP::lam(a, k, j, f) : R(0.0);
        if (f > 0) r -= OP::lam(a, k, j, f - 1);
        return a.rdx * r;
    }
    return a.rdx * (OP::lam(a, k, j, f) - OP::lam(a, k, j, f == 0 ? n - 1 : f - 1));
}

template <class OP>
__device__ real lam_y(const Args& a, int k, int g, int i)
{
    const int n = a.ny;
    if (OP::dupy) {
        real r = (g < n) ? OP::lam(a, k, g, i) : R(0.0);
        if (g > 0) r -= OP::lam(a, k, g - 1, i);
        return a.rdy * r;
    }
    return a.rdy * (OP::lam(a, k, g, i) - OP::lam(a, k, g == 0 ? n - 1 : g - 1, i));
}

template <class OP>
__device__ real field_x(const Args& a, int k, int j, int i)
{
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o) {            // face f reads cell f + o - 3
        const int f = pmod(i + 3 - o, a.nx);
        real v = OP::velx(a, k, j, f);
        acc += (c5(o) * v + s5(o) * fabs(v)) * lam_x<OP>(a, k, j, f);
        if (OP::dupx && f == 0) {
            v = OP::velx(a, k, j, a.nx);
            acc += (c5(o) * v + s5(o) * fabs(v)) * lam_x<OP>(a, k, j, a.nx);
        }
    }
    return acc;
}

template <class OP>
__device__ real field_y(const Args& a, int k, int j, int i)
{
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        const int g = pmod(j + 3 - o, a.ny);
        real v = OP::vely(a, k, g, i);
        acc += (c5(o) * v + s5(o) * fabs(v)) * lam_y<OP>(a, k, g, i);
        if (OP::dupy && g == 0) {
            v = OP::vely(a, k, a.ny, i);
            acc += (c5(o) * v + s5(o) * fabs(v)) * lam_y<OP>(a, k, a.ny, i);
        }
    }
    return acc;
}

// dL/dvel of x face f on the line (k, j)
template <class OP>
__device__ real dvel_x(const Args& a, int k, int j, int f)
{
    const real s = sgn(OP::velx(a, k, j, f));
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o)
        acc += (c5(o) + s5(o) * s) * OP::q(a, k, j, pmod(f + o - 3, a.nx));
    return acc * lam_x<OP>(a, k, j, f);
}

template <class OP>
__device__ real dvel_y(const Args& a, int k, int g, int i)
{
    const real s = sgn(OP::vely(a, k, g, i));
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o)
        acc += (c5(o) + s5(o) * s) * OP::q(a, k, pmod(g + o - 3, a.ny), i);
    return acc * lam_y<OP>(a, k, g, i);
}

// ---------------------------------------------------------------------------
// Vertical direction.  Half-level fields: faces kf = 1..nz-1 at w levels.
// w: faces at mass levels m = 0..nz-1.
// ---------------------------------------------------------------------------

template <class OP>
__device__ real lam_zh(const Args& a, int kf, int j, int i)
{
    return a.rdz[kf] * OP::lam(a, kf, j, i)
           - a.rdz[kf - 1] * OP::lam(a, kf - 1, j, i);
}

template <class OP>
__device__ real field_zh(const Args& a, int k, int j, int i)
{
    const int nz = a.nz;
    real acc = R(0.0);
    for (int kf = k - 1; kf <= k + 2; ++kf) {
        if (kf < 1 || kf > nz - 1) continue;
        const real v = OP::velz(a, kf, j, i);
        const real L = lam_zh<OP>(a, kf, j, i);
        if (kf == 1 || kf == nz - 1) {
            if (k == kf) acc += v * a.fnm[kf] * L;
            else if (k == kf - 1) acc += v * a.fnp[kf] * L;
        } else {
            const int o = k - kf + 2;            // face kf reads kf + o - 2
            acc += (c3(o) * v + s3(o) * fabs(v)) * L;
        }
    }
    return acc;
}

template <class OP>
__device__ real dvel_zh(const Args& a, int kf, int j, int i)
{
    const int nz = a.nz;
    if (kf < 1 || kf > nz - 1) return R(0.0);
    const real L = lam_zh<OP>(a, kf, j, i);
    if (kf == 1 || kf == nz - 1)
        return L * (a.fnm[kf] * OP::q(a, kf, j, i)
                    + a.fnp[kf] * OP::q(a, kf - 1, j, i));
    const real s = sgn(OP::velz(a, kf, j, i));
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 4; ++o)
        acc += (c3(o) + s3(o) * s) * OP::q(a, kf + o - 2, j, i);
    return acc * L;
}

__device__ real lam_zw(const Args& a, int m, int j, int i)
{
    real L = (m >= 1) ? -(a.rdz[m] * OpW::lam(a, m, j, i)) : R(0.0);
    if (m + 1 <= a.nz - 1) L += a.rdz[m + 1] * OpW::lam(a, m + 1, j, i);
    return L;
}

__device__ real field_zw(const Args& a, int k, int j, int i)
{
    const int nz = a.nz;
    real acc = R(0.0);
    for (int m = k - 2; m <= k + 1; ++m) {
        if (m < 0 || m > nz - 1) continue;
        const real v = OpW::velz(a, m, j, i);
        const real L = lam_zw(a, m, j, i);
        if (m == 0 || m == nz - 1) {
            if (k == m || k == m + 1) acc += R(0.5) * v * L;
        } else {
            const int o = k - m + 1;             // face m reads m + o - 1
            acc += (c3(o) * v + s3(o) * fabs(v)) * L;
        }
    }
    return acc;
}

__device__ real dvel_zw(const Args& a, int m, int j, int i)
{
    const int nz = a.nz;
    const real L = lam_zw(a, m, j, i);
    if (m == 0 || m == nz - 1)
        return L * (R(0.5) * (OpW::q(a, m + 1, j, i) + OpW::q(a, m, j, i)));
    const real s = sgn(OpW::velz(a, m, j, i));
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 4; ++o)
        acc += (c3(o) + s3(o) * s) * OpW::q(a, m + o - 1, j, i);
    return acc * L;
}

// ---------------------------------------------------------------------------
// Gathers of the mass-flux gradients: each ru / rv / rw point collects the
// velocity adjoints of the faces whose velocity it enters, with the weight
// it has there.
// ---------------------------------------------------------------------------

template <class OP>
__device__ real grad_ru(const Args& a, int k, int j, int f)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (OP::wlev) {                 // velx(k) = fnm[k] ru[k] + fnp[k] ru[k-1]
        real g = R(0.0);
        if (k >= 1) g += a.fnm[k] * dvel_x<OP>(a, k, j, f);
        if (k + 1 <= nz - 1) g += a.fnp[k + 1] * dvel_x<OP>(a, k + 1, j, f);
        return g;
    }
    if (!OP::dupx) {                // mass centres f-1 and f (faces f, f+1)
        real g = R(0.0);
        if (f < nx) g += dvel_x<OP>(a, k, j, f + 1 == nx ? 0 : f + 1);
        if (f >= 1) g += dvel_x<OP>(a, k, j, f == nx ? 0 : f);
        return R(0.5) * g;
    }
    if (OP::dupy && !OP::wlev)      // scalar: the face itself
        return dvel_x<OP>(a, k, j, f);
    // v: x faces of rows j and j+1 average this row with its neighbour
    return R(0.5) * (dvel_x<OP>(a, k, j, f)
                     + dvel_x<OP>(a, k, j + 1 == ny ? 0 : j + 1, f));
}

template <class OP>
__device__ real grad_rv(const Args& a, int k, int g, int i)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (OP::wlev) {
        real r = R(0.0);
        if (k >= 1) r += a.fnm[k] * dvel_y<OP>(a, k, g, i);
        if (k + 1 <= nz - 1) r += a.fnp[k + 1] * dvel_y<OP>(a, k + 1, g, i);
        return r;
    }
    if (!OP::dupy) {
        real r = R(0.0);
        if (g < ny) r += dvel_y<OP>(a, k, g + 1 == ny ? 0 : g + 1, i);
        if (g >= 1) r += dvel_y<OP>(a, k, g == ny ? 0 : g, i);
        return R(0.5) * r;
    }
    if (OP::dupx)                   // scalar
        return dvel_y<OP>(a, k, g, i);
    // u: y faces of columns i and i+1
    return R(0.5) * (dvel_y<OP>(a, k, g, i)
                     + dvel_y<OP>(a, k, g, i + 1 == nx ? 0 : i + 1));
}

template <class OP>
__device__ real grad_rw(const Args& a, int kf, int j, int i)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    if (OP::wlev) {
        real r = R(0.0);
        if (kf <= nz - 1) r += dvel_zw(a, kf, j, i);
        if (kf >= 1) r += dvel_zw(a, kf - 1, j, i);
        return R(0.5) * r;
    }
    if (OP::dupx && OP::dupy)       // scalar
        return dvel_zh<OP>(a, kf, j, i);
    if (!OP::dupx)                  // u: columns i and i+1
        return R(0.5) * (dvel_zh<OP>(a, kf, j, i)
                         + dvel_zh<OP>(a, kf, j, i + 1 == nx ? 0 : i + 1));
    return R(0.5) * (dvel_zh<OP>(a, kf, j, i)        // v: rows j and j+1
                     + dvel_zh<OP>(a, kf, j + 1 == ny ? 0 : j + 1, i));
}

// One thread per (k, j, i) of the union of the staggers (i <= nx,
// j <= ny, k <= nz); every gradient is added into its output.
template <class OP>
__global__ void k_flux_div_adj(Args a)
{
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    if (a.gq && i < nx && j < ny) {
        if (OP::wlev) {
            real g = field_zw(a, k, j, i);
            if (k >= 1 && k <= nz - 1)
                g += field_x<OP>(a, k, j, i) + field_y<OP>(a, k, j, i);
            a.gq[OP::qidx(a, k, j, i)] += g;
        } else if (k < nz) {
            a.gq[OP::qidx(a, k, j, i)] += field_x<OP>(a, k, j, i)
                                          + field_y<OP>(a, k, j, i)
                                          + field_zh<OP>(a, k, j, i);
        }
    }
    if (a.gru && j < ny && k < nz)
        a.gru[iU(k, j, i, ny, nx)] += grad_ru<OP>(a, k, j, i);
    if (a.grv && i < nx && k < nz)
        a.grv[iV(k, j, i, ny, nx)] += grad_rv<OP>(a, k, j, i);
    if (a.grw && i < nx && j < ny)
        a.grw[iM(k, j, i, ny, nx)] += grad_rw<OP>(a, k, j, i);
}

template <class OP>
int32_t flux_div_adjoint(const real* q, const real* ru, const real* rv,
                         const real* rw, const real* lam, real* g_q,
                         real* g_ru, real* g_rv, real* g_rw, const real* rdz,
                         const real* fnm, const real* fnp, real dx_inv,
                         real dy_inv, int32_t nz, int32_t ny, int32_t nx,
                         void* stream)
{
    if (!q || !ru || !rv || !rw || !lam || !rdz || !fnm || !fnp)
        return lib_fail(ARW_E_ARG, "null argument");
    if (nz < 1 || ny < 1 || nx < 1) return lib_fail(ARW_E_ARG, "bad extents");
    if ((int64_t)(nz + 1) * (ny + 1) * (nx + 1) >= ((int64_t)1 << 31))
        return lib_fail(ARW_E_ARG, "extents exceed 32-bit indexing");
    if (!g_q && !g_ru && !g_rv && !g_rw) return 0;
    Args a{q, ru, rv, rw, lam, rdz, fnm, fnp, g_q, g_ru, g_rv, g_rw,
           dx_inv, dy_inv, nz, ny, nx};
    dim3 grid((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz + 1);
    hipLaunchKernelGGL(k_flux_div_adj<OP>, grid, dim3(64, 4), 0,
                       (hipStream_t)stream, a);
    return lib_launched(hipGetLastError(), "k_flux_div_adj");
}

}  // namespace

}  // namespace ARW_NS

using namespace ARW_NS;

#define ARW_ADV_ADJ(name, OP)                                                \
    extern "C" int32_t ARW_API(name)(                                        \
        const real* q, const real* ru, const real* rv, const real* rw,       \
        const real* lam, real* g_q, real* g_ru, real* g_rv, real* g_rw,      \
        const real* rdz, const real* fnm, const real* fnp, real dx_inv,      \
        real dy_inv, int32_t nz, int32_t ny, int32_t nx, void* stream)       \
    {                                                                        \
        return flux_div_adjoint<OP>(q, ru, rv, rw, lam, g_q, g_ru, g_rv,     \
                                    g_rw, rdz, fnm, fnp, dx_inv, dy_inv, nz, \
                                    ny, nx, stream);                         \
    }

ARW_ADV_ADJ(flux_div_scalar_adjoint, OpScalar)
ARW_ADV_ADJ(flux_div_u_adjoint, OpU)
ARW_ADV_ADJ(flux_div_v_adjoint, OpV)
ARW_ADV_ADJ(flux_div_w_adjoint, OpW)
