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
// atomics and every output is written by exactly one thread; what the
// gathers at one cell share (the output adjoint of its neighbours, the
// velocities of the faces that read it, the field stencil) is loaded once
// per thread.
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
// Horizontal directions.  A line is the x line of (k, j) or the y line of
// (k, i) of one operator: its cells' folded output adjoint and field, and
// the velocity of each face.  "Windows" hold what the gathers at one cell
// need, loaded once: the output adjoint of cells c0-3..c0+3 and the
// velocities of the six faces c0-2..c0+3 that read cell c0.
// ---------------------------------------------------------------------------

// c in [-3, n + 3) to [0, n): one correction when n >= 4
__device__ __forceinline__ int wrapc(int c, int n)
{
    if (n >= 4) return c < 0 ? c + n : (c >= n ? c - n : c);
    return pmod(c, n);
}

template <class OP>
struct XLine {
    const Args& a;
    int k, j;
    static constexpr bool dup = OP::dupx;
    __device__ int n() const { return a.nx; }
    __device__ real rd() const { return a.rdx; }
    __device__ real lam(int c) const { return OP::lam(a, k, j, c); }
    __device__ real q(int c) const { return OP::q(a, k, j, c); }
    __device__ real vel(int f) const { return OP::velx(a, k, j, f); }
};

template <class OP>
struct YLine {
    const Args& a;
    int k, i;
    static constexpr bool dup = OP::dupy;
    __device__ int n() const { return a.ny; }
    __device__ real rd() const { return a.rdy; }
    __device__ real lam(int c) const { return OP::lam(a, k, c, i); }
    __device__ real q(int c) const { return OP::q(a, k, c, i); }
    __device__ real vel(int g) const { return OP::vely(a, k, g, i); }
};

// Face c0-2+t separates cells c0-3+t (left) and c0-2+t (right).  On a dup
// line the periodic face 0 is two faces: face 0 (velocity vel(0), flux into
// the right cell) and face n (vel(n), flux out of the left cell), hence a
// velocity per side.
struct Win {
    real lw[7];           // output adjoint of cells c0-3 .. c0+3
    real vr[6], vl[6];    // velocity of face c0-2+t seen from right / left
};

template <class L>
__device__ __forceinline__ void load_win(const L& l, int c0, Win& w)
{
    const int n = l.n();
#pragma unroll
    for (int p = 0; p < 7; ++p) w.lw[p] = l.lam(wrapc(c0 - 3 + p, n));
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int f = wrapc(c0 - 2 + t, n);
        w.vr[t] = l.vel(f);
        w.vl[t] = (L::dup && f == 0) ? l.vel(n) : w.vr[t];
    }
}

// dL/dq at the window's cell: face c0-2+t reads it as stencil cell
// o = 5 - t, with Lambda = rd (vr lam_right - vl lam_left) split into its
// v and |v| parts.
template <class L>
__device__ __forceinline__ real win_field(const L& l, const Win& w)
{
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        const int t = 5 - o;
        const real lr = w.lw[t + 1], ll = w.lw[t];
        acc += c5(o) * (w.vr[t] * lr - w.vl[t] * ll)
               + s5(o) * (fabs(w.vr[t]) * lr - fabs(w.vl[t]) * ll);
    }
    return l.rd() * acc;
}

// dL/dvel of a face with velocity v, cell adjoints lr / ll on its sides
// (0 where the face has no such cell) and stencil q[f-3 .. f+2].
template <class L>
__device__ __forceinline__ real face_dvel(const L& l, int f, real v, real lr,
                                          real ll)
{
    const int n = l.n();
    const real s = sgn(v);
    real acc = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o)
        acc += (c5(o) + s5(o) * s) * l.q(wrapc(f + o - 3, n));
    return l.rd() * (lr - ll) * acc;
}

// dup line, face f in [0, n], from the window at c0 = f mod n
template <class L>
__device__ __forceinline__ real win_dvel_dup(const L& l, int f, int c0,
                                             const Win& w)
{
    const int n = l.n();
    const real v = (f < n) ? w.vr[2] : w.vl[2];
    const real lr = (f < n) ? w.lw[3] : R(0.0);
    const real ll = (f > 0) ? w.lw[2] : R(0.0);
    return face_dvel(l, c0, v, lr, ll);
}

// dup line, face f in [0, n], without a window (a neighbouring line)
template <class L>
__device__ __forceinline__ real dvel_dup(const L& l, int f)
{
    const int n = l.n();
    const real lr = (f < n) ? l.lam(f) : R(0.0);
    const real ll = (f > 0) ? l.lam(f - 1) : R(0.0);
    return face_dvel(l, f == n ? 0 : f, l.vel(f), lr, ll);
}

// periodic line: dL/dvel of faces c0 and c0+1 from the window at c0, with
// one shared stencil q[c0-3 .. c0+3]
template <class L>
__device__ __forceinline__ void win_dvel_pair(const L& l, int c0,
                                              const Win& w, real& d0,
                                              real& d1)
{
    const int n = l.n();
    real qw[7];
#pragma unroll
    for (int p = 0; p < 7; ++p) qw[p] = l.q(wrapc(c0 - 3 + p, n));
    const real s0 = sgn(w.vr[2]), s1 = sgn(w.vr[3]);
    real a0 = R(0.0), a1 = R(0.0);
#pragma unroll
    for (int o = 0; o < 6; ++o) {
        a0 += (c5(o) + s5(o) * s0) * qw[o];
        a1 += (c5(o) + s5(o) * s1) * qw[o + 1];
    }
    d0 = l.rd() * (w.lw[3] - w.lw[2]) * a0;
    d1 = l.rd() * (w.lw[4] - w.lw[3]) * a1;
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
// The kernels: one thread per (k, j, i) of the union of the staggers
// (i <= nx, j <= ny, k <= nz).  A thread gathers the field gradient of its
// cell and the mass-flux gradients of its ru / rv / rw points; each mass
// flux collects the face velocity adjoints of the faces whose velocity it
// enters, with its weight there.  Every gradient is added into its output.
// ---------------------------------------------------------------------------

// scalar: faces are the mass-flux points themselves
__global__ void k_adj_scalar(Args a)
{
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const int ic = (i == nx) ? 0 : i, jc = (j == ny) ? 0 : j;
    real gq = R(0.0);
    if (k < nz && j < ny) {
        XLine<OpScalar> l{a, k, j};
        Win w;
        load_win(l, ic, w);
        if (i < nx) gq += win_field(l, w);
        if (a.gru) a.gru[iU(k, j, i, ny, nx)] += win_dvel_dup(l, i, ic, w);
    }
    if (k < nz && i < nx) {
        YLine<OpScalar> l{a, k, i};
        Win w;
        load_win(l, jc, w);
        if (j < ny) gq += win_field(l, w);
        if (a.grv) a.grv[iV(k, j, i, ny, nx)] += win_dvel_dup(l, j, jc, w);
    }
    if (i < nx && j < ny) {
        if (a.gq && k < nz)
            a.gq[iM(k, j, i, ny, nx)] += gq + field_zh<OpScalar>(a, k, j, i);
        if (a.grw) a.grw[iM(k, j, i, ny, nx)] += dvel_zh<OpScalar>(a, k, j, i);
    }
}

// u: x faces at mass centres (periodic, velocity 0.5 (ru[m] + ru[m+1]));
// y and z faces average rv / rw over the columns i-1 and i
__global__ void k_adj_u(Args a)
{
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const int ic = (i == nx) ? 0 : i, jc = (j == ny) ? 0 : j;
    const int ip = (i + 1 >= nx) ? i + 1 - nx : i + 1;
    real gq = R(0.0);
    if (k < nz && j < ny) {
        XLine<OpU> l{a, k, j};
        Win w;
        load_win(l, ic, w);
        if (i < nx) gq += win_field(l, w);
        if (a.gru) {           // ru[i] enters mass centres i-1 and i
            real d0, d1;
            win_dvel_pair(l, ic, w, d0, d1);
            a.gru[iU(k, j, i, ny, nx)] +=
                R(0.5) * ((i < nx ? d1 : R(0.0)) + (i >= 1 ? d0 : R(0.0)));
        }
    }
    if (k < nz && i < nx) {
        YLine<OpU> l{a, k, i};
        Win w;
        load_win(l, jc, w);
        if (j < ny) gq += win_field(l, w);
        if (a.grv)             // rv[., i] enters the y faces of u columns i, i+1
            a.grv[iV(k, j, i, ny, nx)] +=
                R(0.5) * (win_dvel_dup(l, j, jc, w)
                          + dvel_dup(YLine<OpU>{a, k, ip}, j));
    }
    if (i < nx && j < ny) {
        if (a.gq && k < nz)
            a.gq[iU(k, j, i, ny, nx)] += gq + field_zh<OpU>(a, k, j, i);
        if (a.grw)
            a.grw[iM(k, j, i, ny, nx)] +=
                R(0.5) * (dvel_zh<OpU>(a, k, j, i) + dvel_zh<OpU>(a, k, j, ip));
    }
}

// v: the mirror of u
__global__ void k_adj_v(Args a)
{
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const int ic = (i == nx) ? 0 : i, jc = (j == ny) ? 0 : j;
    const int jp = (j + 1 >= ny) ? j + 1 - ny : j + 1;
    real gq = R(0.0);
    if (k < nz && j < ny) {
        XLine<OpV> l{a, k, j};
        Win w;
        load_win(l, ic, w);
        if (i < nx) gq += win_field(l, w);
        if (a.gru)             // ru[j, .] enters the x faces of v rows j, j+1
            a.gru[iU(k, j, i, ny, nx)] +=
                R(0.5) * (win_dvel_dup(l, i, ic, w)
                          + dvel_dup(XLine<OpV>{a, k, jp}, i));
    }
    if (k < nz && i < nx) {
        YLine<OpV> l{a, k, i};
        Win w;
        load_win(l, jc, w);
        if (j < ny) gq += win_field(l, w);
        if (a.grv) {           // rv[j] enters mass centres j-1 and j
            real d0, d1;
            win_dvel_pair(l, jc, w, d0, d1);
            a.grv[iV(k, j, i, ny, nx)] +=
                R(0.5) * ((j < ny ? d1 : R(0.0)) + (j >= 1 ? d0 : R(0.0)));
        }
    }
    if (i < nx && j < ny) {
        if (a.gq && k < nz)
            a.gq[iV(k, j, i, ny, nx)] += gq + field_zh<OpV>(a, k, j, i);
        if (a.grw)
            a.grw[iM(k, j, i, ny, nx)] +=
                R(0.5) * (dvel_zh<OpV>(a, k, j, i) + dvel_zh<OpV>(a, k, jp, i));
    }
}

// w: horizontal faces at levels 1..nz-1 with velocities fnm[k] r(k) +
// fnp[k] r(k-1); vertical faces at mass levels
__global__ void k_adj_w(Args a)
{
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const int ic = (i == nx) ? 0 : i, jc = (j == ny) ? 0 : j;
    const bool own = (k >= 1 && k <= nz - 1);      // level k has faces
    const bool up = (k + 1 <= nz - 1);             // so has level k+1
    real gq = R(0.0);
    if (j < ny && k < nz) {
        real g = R(0.0);
        if (own) {
            XLine<OpW> l{a, k, j};
            Win w;
            load_win(l, ic, w);
            if (i < nx) gq += win_field(l, w);
            g = a.fnm[k] * win_dvel_dup(l, i, ic, w);
        }
        if (a.gru) {           // ru[k] enters levels k (fnm) and k+1 (fnp)
            if (up) g += a.fnp[k + 1] * dvel_dup(XLine<OpW>{a, k + 1, j}, i);
            a.gru[iU(k, j, i, ny, nx)] += g;
        }
    }
    if (i < nx && k < nz) {
        real g = R(0.0);
        if (own) {
            YLine<OpW> l{a, k, i};
            Win w;
            load_win(l, jc, w);
            if (j < ny) gq += win_field(l, w);
            g = a.fnm[k] * win_dvel_dup(l, j, jc, w);
        }
        if (a.grv) {
            if (up) g += a.fnp[k + 1] * dvel_dup(YLine<OpW>{a, k + 1, i}, j);
            a.grv[iV(k, j, i, ny, nx)] += g;
        }
    }
    if (i < nx && j < ny) {
        if (a.gq) a.gq[iM(k, j, i, ny, nx)] += gq + field_zw(a, k, j, i);
        if (a.grw) {           // rw[kf] enters the faces at m = kf-1 and kf
            real r = R(0.0);
            if (k <= nz - 1) r += dvel_zw(a, k, j, i);
            if (k >= 1) r += dvel_zw(a, k - 1, j, i);
            a.grw[iM(k, j, i, ny, nx)] += R(0.5) * r;
        }
    }
}

void launch_adj(OpScalar, dim3 g, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL(k_adj_scalar, g, dim3(64, 4), 0, s, a);
}
void launch_adj(OpU, dim3 g, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL(k_adj_u, g, dim3(64, 4), 0, s, a);
}
void launch_adj(OpV, dim3 g, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL(k_adj_v, g, dim3(64, 4), 0, s, a);
}
void launch_adj(OpW, dim3 g, hipStream_t s, const Args& a)
{
    hipLaunchKernelGGL(k_adj_w, g, dim3(64, 4), 0, s, a);
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
    launch_adj(OP{}, grid, (hipStream_t)stream, a);
    return lib_launched(hipGetLastError(), "k_adj (advection adjoint)");
}

}  // namespace

int32_t lib_flux_div_adjoint(int op, const real* q, const real* ru,
                             const real* rv, const real* rw, const real* lam,
                             real* g_q, real* g_ru, real* g_rv, real* g_rw,
                             const real* rdz, const real* fnm,
                             const real* fnp, real dx_inv, real dy_inv,
                             int nz, int ny, int nx, hipStream_t st)
{
    switch (op) {
    case 0:
        return flux_div_adjoint<OpScalar>(q, ru, rv, rw, lam, g_q, g_ru, g_rv,
                                          g_rw, rdz, fnm, fnp, dx_inv, dy_inv,
                                          nz, ny, nx, st);
    case 1:
        return flux_div_adjoint<OpU>(q, ru, rv, rw, lam, g_q, g_ru, g_rv, g_rw,
                                     rdz, fnm, fnp, dx_inv, dy_inv, nz, ny,
                                     nx, st);
    case 2:
        return flux_div_adjoint<OpV>(q, ru, rv, rw, lam, g_q, g_ru, g_rv, g_rw,
                                     rdz, fnm, fnp, dx_inv, dy_inv, nz, ny,
                                     nx, st);
    default:
        return flux_div_adjoint<OpW>(q, ru, rv, rw, lam, g_q, g_ru, g_rv, g_rw,
                                     rdz, fnm, fnp, dx_inv, dy_inv, nz, ny,
                                     nx, st);
    }
}

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
