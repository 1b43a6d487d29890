// Hand-written adjoint of one acoustic substep (k_uv + k_column, the
// reference's advance_uv -> advance_mu_th -> advance_w_phi, acoustic.cu:
// 160-982).
//
// For a frozen stage reference state the substep maps the small-step
// state X = (u'', v'', w'', theta'', phi'', mu'', p'', p''_old, alpha'') and
// the slow tendencies to the new X.  It is affine in X except for the new
// column mass mu_t + mu'' in three denominators (the theta'' term of the w
// equation, the phi'' update, the linearised EOS), so J depends on X: the
// adjoint first re-runs the forward substep on copies of X (the same
// kernels as arw_acoustic_substep) and then applies J^T at X.
//
// Passes:
//   k_adj_p1/s1/p2/p3/s2  mass columns: the vertical reverse of the w/phi
//                solve, the EOS, the Thomas solve, the theta'' and
//                mu''/Omega'' updates, as level-parallel kernels around two
//                column-serial ones; leaves per-cell adjoints for the
//                horizontal parts
//   k_adj_f1/f2  u and v faces, per level: adjoint of u''/v'' (gathered
//                from the two cells of each face) and the per-face pieces
//                of the advance_uv transpose
//   k_adj_cells  mass columns: gather of the face pieces into p'', p''_old,
//                phi'', mu'', alpha''
#include "common.h"
#include "march.h"

#include <math.h>

namespace ARW_NS {

namespace {

// work slots (each (nz+1)*(ny+1)*(nx+1) elements)
enum {
    W_U, W_V, W_W, W_TH, W_PH, W_MU, W_P, W_PO, W_AL, W_WW,   // forward copies
    W_LPH1, W_LW, W_SA, W_SG, W_LRHS, W_LT2, W_LWD, W_LWW,   // column scratch
    W_GG, W_LDVP, W_L0,                                      // cell -> faces
    W_MC1, W_MC2, W_MC3, W_GGF, W_G0,                        // level pieces
    W_LBX, W_LSX, W_LBY, W_LSY,                              // faces -> cells
    W_COUNT
};

struct Adj {
    const real *lu, *lv, *lw, *lth, *lph, *lmu, *lp, *lpo, *lal, *lww;
    real *gu, *gv, *gw, *gth, *gph, *gmu, *gp, *gpo, *gal;
    real *gru, *grv, *grw, *grth, *grph, *grmu;
};

struct Fw {                 // forward recompute and scratch
    const real *M1, *T1, *W1, *AL1;
    real* s[W_COUNT];
};

__device__ __forceinline__ real th_ref(const Dom& d, int k, int c)
{
    return (d.base3d ? d.thb[k * d.ny * d.nx + c] : d.thb[k]) + d.thp[k * d.ny * d.nx + c];
}

__device__ __forceinline__ real phi_full(const Dom& d, int k, int c)
{
    return (d.base3d ? d.phb[k * d.ny * d.nx + c] : d.phb[k]) + d.php[k * d.ny * d.nx + c];
}

// The mass-column reverse of advance_w_phi + diagnose_p and of the
// vertical parts of advance_mu_th, split by dependence: level-parallel
// kernels (one thread per mass point and level, recomputing the few
// neighbour-level terms they need) around two lean column-serial kernels
// (the transposed Thomas solve; the column sums and the Omega'' suffix
// sums).  Per-level pieces of column sums go to MC1..MC3.
struct ColCtx {
    const Dom& d;
    const Sc& s;
    int c, st, nz;
    real mut, M1, K, cof;
    __device__ ColCtx(const Dom& dd, const Sc& ss, int cc, real m1)
        : d(dd), s(ss), c(cc), st(dd.ny * dd.nx), nz(dd.nz)
    {
        mut = d.mub2d[c] + d.mup[c];
        M1 = m1;
        K = R(0.5) * s.dtau * G_F * (R(1.0) + s.epssm);
        cof = K * K;
    }
    __device__ real c2a(int k) const
    {
        return (GAMMA_F * d.p[k * st + c]) / d.alt[k * st + c];
    }
    __device__ real chm(int k) const { return d.c1h[k] * mut + d.c2h[k]; }
    __device__ real cfm(int k) const { return d.c1f[k] * mut + d.c2f[k]; }
    __device__ real chmp(int k) const { return d.c1h[k] * (mut + M1) + d.c2h[k]; }
    __device__ real cfs(int k) const { return d.c1f[k] * (mut + M1) + d.c2f[k]; }
    __device__ real a(int k) const
    {
        if (k >= 2 && k < nz)
            return -cof * d.rdn[k] * d.rdnw[k - 1] * c2a(k - 1)
                   / (chm(k - 1) * cfm(k - 1));
        if (k == nz && !s.top_lid)
            return R(-2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1] * c2a(nz - 1)
                   / (chm(nz - 1) * cfm(nz - 1));
        return R(0.0);
    }
    // 1/(1+dw) of the damp_opt = 3 damper at w level k (1 where undamped)
    __device__ real undamp(int k) const
    {
        if (!s.damp) return R(1.0);
        const real g = G_F;
        const real hbot = phi_full(d, nz, c) / g - s.zdamp;
        const real hk = phi_full(d, k, c) / g;
        if (!(hk >= hbot)) return R(1.0);
        const real arg = R(1.5707963) * (hk - hbot) / s.zdamp;
        const real sn = (real)sin((double)arg);
        return R(1.0) / (R(1.0) + s.dampmag * sn * sn);
    }
    // lambda(alpha'') of level k from the EOS reverse
    __device__ real lAL(const Adj& A, int k) const
    {
        const int x = k * st + c;
        return A.lal[x] - c2a(k) * A.lp[x];
    }
    // lambda of w row k (rows 1..nz-1, and nz with an open lid)
    __device__ bool row(int k) const
    {
        return (k >= 1 && k < nz) || (k == nz && !s.top_lid);
    }
    __device__ real lup(int k, real la) const
    {
        if (k < 1 || k >= nz) return R(0.0);
        return R(0.5) * s.dtau * G_F * d.rdn[k] * c2a(k) * d.rdnw[k] / chm(k) * la;
    }
    __device__ real ldn(int k, real la) const
    {
        if (k >= 1 && k < nz)
            return -(R(0.5) * s.dtau * G_F * d.rdn[k] * c2a(k - 1) * d.rdnw[k - 1]
                     / chm(k - 1)) * la;
        if (k == nz && !s.top_lid)
            return -(s.dtau * G_F * c2a(nz - 1) * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                     / chm(nz - 1)) * la;
        return R(0.0);
    }
};

__device__ __forceinline__ bool col_level(const Dom& d, int& c, int& k)
{
    c = blockIdx.x * blockDim.x + threadIdx.x;
    k = blockIdx.y;
    return c < d.ny * d.nx;
}

// P1: EOS reverse, phi'' update reverse, damper; w-solve right-hand side.
__global__ __launch_bounds__(256) void k_adj_p1(Dom d, Sc s, Adj a, Fw f)
{
    int c, k;
    if (!col_level(d, c, k)) return;
    const ColCtx C(d, s, c, f.M1[c]);
    const int st = C.st, nz = d.nz, x = k * st + c;
    real* LRHS = f.s[W_LRHS];
    real* LWIN = f.s[W_LPH1];
    real mcon = R(0.0);
    real ph1 = a.lph[x];
    if (k < nz) {
        const real thr = th_ref(d, k, c), alt = d.alt[x], c2a = C.c2a(k);
        const real chmp = C.chmp(k);
        const real q = alt / (chmp * thr);
        const real X = f.T1[x] - d.c1h[k] * C.M1 * thr;
        const real lPP = a.lp[x];
        f.s[W_GG][x] = a.lth[x] + c2a * q * lPP;
        mcon += c2a * (-q * d.c1h[k] * thr - q * X * d.c1h[k] / chmp) * lPP;
        const real lAL = a.lal[x] - c2a * lPP;
        mcon += (-(alt * d.c1h[k]) / chmp - f.AL1[x] * d.c1h[k] / chmp) * lAL;
        ph1 += d.rdnw[k] / chmp * lAL;
    }
    if (k >= 1) ph1 += -d.rdnw[k - 1] / C.chmp(k - 1) * C.lAL(a, k - 1);
    if (k == 0) {
        f.s[W_G0][c] = ph1;                    // phi''(0) passes through
        LWIN[c] = a.lw[c];
    } else {
        LRHS[x] = ph1;
        const real cfs = C.cfs(k);
        mcon += -C.K * f.W1[x] * d.c1f[k] / (cfs * cfs) * ph1;
        const real lw = a.lw[x] + C.K / cfs * ph1;
        LWIN[x] = s.damp ? lw * C.undamp(k) : lw;
    }
    f.s[W_MC1][x] = mcon;
}

// S1: the transposed Thomas solve of one column.
__global__ __launch_bounds__(64) void k_adj_s1(Dom d, Sc s, Fw f)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ny * d.nx) return;
    const ColCtx C(d, s, c, f.M1[c]);
    const int st = C.st, nz = d.nz;
    const real one = R(1.0);
    const real* __restrict__ LWIN = f.s[W_LPH1];
    real* __restrict__ LW = f.s[W_LW];
    real* __restrict__ SA = f.s[W_SA];
    real* __restrict__ SG = f.s[W_SG];
    real gprev = R(0.0), vprev = R(0.0);
    real c2a_k = C.c2a(0), c2a_m;
    for (int k = 1; k < nz; ++k) {
        c2a_m = c2a_k;
        c2a_k = C.c2a(k);
        const real b = one + C.cof * d.rdn[k] * (
            d.rdnw[k] * c2a_k / (C.chm(k) * C.cfm(k))
            + d.rdnw[k - 1] * c2a_m / (C.chm(k - 1) * C.cfm(k)));
        const real cc = -C.cof * d.rdn[k] * d.rdnw[k] * c2a_k
                        / (C.chm(k) * C.cfm(k + 1));
        const real ak = C.a(k);
        const real al = one / (b - ak * gprev);
        const real v = LWIN[k * st + c] - gprev * vprev;
        gprev = cc * al;
        vprev = v;
        LW[k * st + c] = v;
        SA[k * st + c] = al;
        SG[k * st + c] = ak;
    }
    {
        const real bn = one + R(2.0) * C.cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                        * c2a_k / (C.chm(nz - 1) * C.cfm(nz));
        const real an = C.a(nz);
        LW[nz * st + c] = LWIN[nz * st + c] - gprev * vprev;
        SA[nz * st + c] = one / (bn - an * gprev);
        SG[nz * st + c] = an;
    }
    real carry = R(0.0);
    for (int k = nz; k >= 1; --k) {
        const int x = k * st + c;
        const real v = SA[x] * (LW[x] + carry);
        LW[x] = v;
        carry = -SG[x] * v;
    }
    f.s[W_L0][c] = LWIN[c] + carry;
}

// P2: reverse of the w accumulation.
__global__ __launch_bounds__(256) void k_adj_p2(Dom d, Sc s, Adj a, Fw f)
{
    int c, k;
    if (!col_level(d, c, k)) return;
    const ColCtx C(d, s, c, f.M1[c]);
    const int st = C.st, nz = d.nz, x = k * st + c;
    const real one = R(1.0), dt = s.dtau, g = G_F, eps = s.epssm;
    const real* LW = f.s[W_LW];
    auto la = [&](int kk) { return C.row(kk) ? LW[kk * st + c] : R(0.0); };
    const real la_k = la(k), la_kp = la(k + 1);
    const real T = (C.lup(k - 1, la(k - 1)) - C.lup(k, la_k))
                   + (C.ldn(k, la_k) - C.ldn(k + 1, la_kp));
    f.s[W_LWD][x] = (one - eps) * T;                      // GPH
    if (k >= 1) f.s[W_LRHS][x] = f.s[W_LRHS][x] + (one + eps) * T;
    const bool row = C.row(k);
    a.gw[x] = row ? la_k : R(0.0);
    a.grw[x] = row ? dt * la_k : R(0.0);
    f.s[W_MC2][x] = row ? -dt * g * d.c1f[k] * la_k : R(0.0);
    if (k < nz) {
        real t2 = R(0.0);
        if (k >= 1) t2 += dt * g * d.rdn[k] * C.c2a(k) * d.alt[x] * la_k;
        if (k + 1 < nz)
            t2 -= dt * g * d.rdn[k + 1] * C.c2a(k) * d.alt[x] * la_kp;
        if (k == nz - 1 && !s.top_lid)
            t2 -= dt * g * R(2.0) * d.rdnw[nz - 1] * C.c2a(nz - 1) * d.alt[x] * la_kp;
        f.s[W_LT2][x] = t2;
    }
}

// P3: rhs reverse (phi'', rph_t, w''), theta'' adjoint, Omega'' adjoint.
__global__ __launch_bounds__(256) void k_adj_p3(Dom d, Sc s, Adj a, Fw f)
{
    int c, k;
    if (!col_level(d, c, k)) return;
    const ColCtx C(d, s, c, f.M1[c]);
    const int st = C.st, nz = d.nz, x = k * st + c;
    const real half = R(0.5), one = R(1.0), dt = s.dtau, g = G_F, eps = s.epssm;
    const real* LRHS = f.s[W_LRHS];
    const real* LT2 = f.s[W_LT2];
    auto lr2 = [&](int kk) -> real {
        if (kk < 1 || kk > nz || (kk == nz && s.top_lid)) return R(0.0);
        return LRHS[kk * st + c] / C.cfm(kk);
    };
    auto lrhs = [&](int kk) -> real {
        if (kk < 1 || kk > nz || (kk == nz && s.top_lid)) return R(0.0);
        return LRHS[kk * st + c];
    };
    auto lwd = [&](int kk) -> real {            // lambda(wd[kk]), kk < nz
        real v = R(0.0);
        if (kk >= 1) v += -dt * d.fnm[kk] * lr2(kk);
        if (kk + 1 <= nz - 1) v += -dt * d.fnp[kk + 1] * lr2(kk + 1);
        return v;
    };
    auto cwd = [&](int kk) {
        return half * d.rdnw[kk] * (phi_full(d, kk + 1, c) - phi_full(d, kk, c));
    };
    auto lT1 = [&](int kk) {
        const real chs = d.c1h[kk] * (C.mut + C.M1) + d.c2h[kk];
        return f.s[W_GG][kk * st + c]
               + half * (one + eps) / (chs * th_ref(d, kk, c)) * LT2[kk * st + c];
    };
    a.gph[x] = f.s[W_LWD][x] + (k == 0 ? f.s[W_G0][c] : lrhs(k));
    a.grph[x] = k >= 1 ? dt * lr2(k) : R(0.0);
    if (k >= 1) a.gw[x] = a.gw[x] + dt * half * g * (one - eps) * lr2(k);
    real ww = a.lww ? a.lww[x] : R(0.0);
    if (k < nz) ww += cwd(k) * lwd(k);
    if (k >= 1) ww += cwd(k - 1) * lwd(k - 1);
    if (k < nz) {
        const real thr = th_ref(d, k, c);
        const real chs = d.c1h[k] * (C.mut + C.M1) + d.c2h[k];
        const real t2 = half * ((one + eps) * f.T1[x] + (one - eps) * d.th_pp[x])
                        / (chs * thr);
        const real lt2 = LT2[x];
        const real l1 = lT1(k);
        a.gth[x] = half * (one - eps) / (chs * thr) * lt2 + l1;
        a.grth[x] = dt * l1;
        f.s[W_GGF][x] = -dt * l1;
        f.s[W_MC3][x] = -t2 * d.c1h[k] / chs * lt2;
        if (k >= 1) {
            const real lwdtn = d.rdnw[k - 1] * (-dt * lT1(k - 1))
                               - d.rdnw[k] * (-dt * l1);
            ww += (d.fnm[k] * thr + d.fnp[k] * th_ref(d, k - 1, c)) * lwdtn;
        }
    }
    f.s[W_LWW][x] = ww;
}

// S2: column sums of the mu'' pieces, Omega'' suffix sums, and the
// divergence adjoint.
__global__ __launch_bounds__(64) void k_adj_s2(Dom d, Sc s, Adj a, Fw f)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ny * d.nx) return;
    const int st = d.ny * d.nx, nz = d.nz;
    const real half = R(0.5), one = R(1.0), dt = s.dtau, eps = s.epssm;
    const real* __restrict__ MC1 = f.s[W_MC1];
    const real* __restrict__ MC2 = f.s[W_MC2];
    const real* __restrict__ MC3 = f.s[W_MC3];
    const real* __restrict__ LWW = f.s[W_LWW];
    real* __restrict__ LDVP = f.s[W_LDVP];
    real lM1 = a.lmu[c], lmuave = R(0.0);
    for (int k = 0; k <= nz; ++k) {
        lM1 += MC1[k * st + c];
        lmuave += MC2[k * st + c];
        if (k < nz) lM1 += MC3[k * st + c];
    }
    const real lM0old = half * (one - eps) * lmuave;
    lM1 += half * (one + eps) * lmuave;
    real acc = R(0.0), ldmdt = R(0.0), gr = R(0.0);
    for (int k = nz - 1; k >= 0; --k) {
        real linc = R(0.0);
        if (k < nz - 1) {
            acc += LWW[(k + 1) * st + c];
            linc = -acc;
        }
        LDVP[k * st + c] = d.dnw[k] * linc;
        ldmdt += d.dnw[k] * d.c1h[k] * linc;
        gr += d.dnw[k] * d.c1h[k] * linc;
    }
    ldmdt += dt * lM1;
    gr += dt * lM1;
    a.grmu[c] = gr;
    a.gmu[c] = lM1 + lM0old;
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        LDVP[x] = LDVP[x] + d.dnw[k] * ldmdt;
    }
}

// Face pass: u faces (j, fc), fc = 0..nx, then v faces (g, i), g = 0..ny.
// lambda of u''_new (v''_new) is the output adjoint plus what the mass
// cells on both sides took from it (divergence, theta flux, lower
// boundary).  For the non-duplicate faces also the advance_uv transpose
// pieces: the adjoint of the bracket (LB), of the face-averaged damped p''
// levels folded onto the p'' levels (LS), and of the face mass sum (LM).
struct FaceGeo {
    int out0, ostride;       // output adjoint: index of level 0, level stride
    int cE, cEp;             // cell whose east/north face this is, its +1 cell
    int cW, cWm;             // cell whose west/south face this is, its -1 cell
    bool hasE, hasW, own;
    int cA, cB, slot;        // advance_uv cells of an own face, table slot
};

// Geometry of the face of thread t (u faces first, then v faces); dg is
// the duplicate face (nx, or ny) whose adjoint face 0 folds in.
__device__ __forceinline__ bool face_geo(const Dom& d, int t, FaceGeo& fg,
                                         FaceGeo& dg, bool& has_dup,
                                         bool& isu)
{
    const int ny = d.ny, nx = d.nx;
    const int nu = ny * (nx + 1), nv = (ny + 1) * nx;
    if (t >= nu + nv) return false;
    has_dup = false;
    isu = t < nu;
    if (isu) {
        const int j = t / (nx + 1), fc = t - j * (nx + 1);
        auto geo = [&](int f1, FaceGeo& q) {
            q.out0 = iU(0, j, f1, ny, nx);
            q.ostride = ny * (nx + 1);
            q.hasE = f1 >= 1;
            q.cE = j * nx + (f1 >= 1 ? f1 - 1 : 0);
            q.cEp = j * nx + (f1 == nx ? 0 : f1);
            q.hasW = f1 < nx;
            q.cW = j * nx + (f1 < nx ? f1 : 0);
            q.cWm = j * nx + (f1 == 0 ? nx - 1 : (f1 < nx ? f1 - 1 : 0));
            q.own = f1 < nx;
            q.cA = j * nx + (f1 < nx ? f1 : 0);
            q.cB = j * nx + (f1 == 0 || f1 == nx ? nx - 1 : f1 - 1);
            q.slot = j * nx + (f1 < nx ? f1 : 0);
        };
        geo(fc, fg);
        if (fc == 0) { geo(nx, dg); has_dup = true; }
    } else {
        const int r = t - nu, g = r / nx, i = r - g * nx;
        auto geo = [&](int g1, FaceGeo& q) {
            q.out0 = iV(0, g1, i, ny, nx);
            q.ostride = (ny + 1) * nx;
            q.hasE = g1 >= 1;
            q.cE = (g1 >= 1 ? g1 - 1 : 0) * nx + i;
            q.cEp = (g1 == ny ? 0 : g1) * nx + i;
            q.hasW = g1 < ny;
            q.cW = (g1 < ny ? g1 : 0) * nx + i;
            q.cWm = (g1 == 0 ? ny - 1 : (g1 < ny ? g1 - 1 : 0)) * nx + i;
            q.own = g1 < ny;
            q.cA = (g1 < ny ? g1 : 0) * nx + i;
            q.cB = (g1 == 0 || g1 == ny ? ny - 1 : g1 - 1) * nx + i;
            q.slot = (g1 < ny ? g1 : 0) * nx + i;
        };
        geo(g, fg);
        if (g == 0) { geo(ny, dg); has_dup = true; }
    }
    return true;
}

// F1 (face, level): lambda of u''_new / v''_new -> the u''/v'' and ru_t/rv_t
// adjoints, and for own faces the bracket (LB) and dpn-level (LT) pieces.
__global__ __launch_bounds__(256) void k_adj_f1(Dom d, Sc s, Adj a, Fw f)
{
    FaceGeo fg, dg;
    bool has_dup, isu;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (!face_geo(d, t, fg, dg, has_dup, isu)) return;
    const int st = d.ny * d.nx;
    const real half = R(0.5), dt = s.dtau;
    const real rd = isu ? s.rdx : s.rdy;
    const real* lout = isu ? a.lu : a.lv;
    const real* GG = f.s[W_GGF];
    const real* LDVP = f.s[W_LDVP];
    const real* L0 = f.s[W_L0];
    const real cf[3] = {d.cf1, d.cf2, d.cf3};
    auto lam1 = [&](const FaceGeo& q) -> real {
        real v = lout[q.out0 + k * q.ostride];
        if (q.hasE) {
            const int c = q.cE, cp = q.cEp;
            v += half * rd * (th_ref(d, k, cp) + th_ref(d, k, c)) * GG[k * st + c];
            v += rd * LDVP[k * st + c];
            if (k < 3) v += cf[k] * half * rd * (d.ht[cp] - d.ht[c]) * L0[c];
        }
        if (q.hasW) {
            const int c = q.cW, cm = q.cWm;
            v -= half * rd * (th_ref(d, k, c) + th_ref(d, k, cm)) * GG[k * st + c];
            v -= rd * LDVP[k * st + c];
            if (k < 3) v += cf[k] * half * rd * (d.ht[c] - d.ht[cm]) * L0[c];
        }
        return v;
    };
    const real l1 = lam1(fg);
    (isu ? a.gu : a.gv)[fg.out0 + k * fg.ostride] = l1;
    (isu ? a.gru : a.grv)[fg.out0 + k * fg.ostride] = dt * l1;
    if (!fg.own) return;
    const real ldup = has_dup ? lam1(dg) : R(0.0);
    const real ld = -dt * l1 - dt * ldup;
    const real mA = d.mub2d[fg.cA] + d.mup[fg.cA];
    const real mB = d.mub2d[fg.cB] + d.mup[fg.cB];
    const real muf = half * (mB + mA);
    const int x = k * st + fg.slot;
    f.s[isu ? W_LBX : W_LBY][x] = half * rd * (d.c1h[k] * muf + d.c2h[k]) * ld;
    real dphp = half * ((d.php[k * st + fg.cA] + d.php[(k + 1) * st + fg.cA])
                        - (d.php[k * st + fg.cB] + d.php[(k + 1) * st + fg.cB]));
    if (d.base3d)
        dphp = dphp + half * ((d.phb[k * st + fg.cA] + d.phb[(k + 1) * st + fg.cA])
                              - (d.phb[k * st + fg.cB] + d.phb[(k + 1) * st + fg.cB]));
    f.s[isu ? W_LPH1 : W_LW][x] = rd * dphp * ld;
}

// F2 (own face, level): the damped-p'' level adjoints folded onto the p''
// levels (LS), from LT at the neighbouring levels.
__global__ __launch_bounds__(256) void k_adj_f2(Dom d, Sc s, Fw f)
{
    FaceGeo fg, dg;
    bool has_dup, isu;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (!face_geo(d, t, fg, dg, has_dup, isu) || !fg.own) return;
    const int st = d.ny * d.nx, nz = d.nz;
    const real half = R(0.5);
    const real cf[3] = {d.cf1, d.cf2, d.cf3};
    const real* LT = f.s[isu ? W_LPH1 : W_LW];
    auto ldpn = [&](int kf) -> real {
        real v = R(0.0);
        if (kf >= 1) v += d.rdnw[kf - 1] * LT[(kf - 1) * st + fg.slot];
        if (kf <= nz - 1) v -= d.rdnw[kf] * LT[kf * st + fg.slot];
        return v;
    };
    real v = R(0.0);
    if (k >= 1) v += half * d.fnm[k] * ldpn(k);
    if (k + 1 <= nz - 1) v += half * d.fnp[k + 1] * ldpn(k + 1);
    if (k < 3) v += half * cf[k] * ldpn(0);
    if (s.top_lid && k >= nz - 3) v += half * cf[nz - 1 - k] * ldpn(nz);
    f.s[isu ? W_LSX : W_LSY][k * st + fg.slot] = v;
}

// Cell pass: gather the face pieces of the advance_uv transpose.
__global__ void k_adj_cells(Dom d, Sc s, Adj a, Fw f)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = j * nx + (i + 1 == nx ? 0 : i + 1);
    const int im = j * nx + (i == 0 ? nx - 1 : i - 1);
    const int jp = (j + 1 == ny ? 0 : j + 1) * nx + i;
    const int jm = (j == 0 ? ny - 1 : j - 1) * nx + i;
    const real* LBX = f.s[W_LBX];
    const real* LSX = f.s[W_LSX];
    const real* LBY = f.s[W_LBY];
    const real* LSY = f.s[W_LSY];
    const real one = R(1.0), sm = s.smdiv;
    const real* LTX = f.s[W_LPH1];
    const real* LTY = f.s[W_LW];
    real lms = R(0.0);
    for (int k = 0; k < nz; ++k) {
        const int x = k * st;
        // the face mass sums: -1/2 c1h t of the faces the cell touches
        lms += -R(0.5) * d.c1h[k] * ((LTX[x + c] + LTX[x + ip])
                                     + (LTY[x + c] + LTY[x + jp]));
        auto alt = [&](int cc) { return d.alt[x + cc]; };
        // face c is the cell's own (A side), face ip/jp has it as B side
        real lpd = (alt(c) + alt(im)) * LBX[x + c] - (alt(ip) + alt(c)) * LBX[x + ip]
                   + (alt(c) + alt(jm)) * LBY[x + c] - (alt(jp) + alt(c)) * LBY[x + jp];
        lpd += LSX[x + c] + LSX[x + ip] + LSY[x + c] + LSY[x + jp];
        const real lb = (LBX[x + c] - LBX[x + ip]) + (LBY[x + c] - LBY[x + jp]);
        a.gph[x + c] += lb;
        a.gph[x + st + c] += lb;
        real gal = R(0.0);
        if (d.base3d) {
            auto pb = [&](int cc) { return d.pb[x + cc]; };
            gal = (pb(c) - pb(im)) * LBX[x + c] + (pb(ip) - pb(c)) * LBX[x + ip]
                  + (pb(c) - pb(jm)) * LBY[x + c] + (pb(jp) - pb(c)) * LBY[x + jp];
        }
        a.gal[x + c] = gal;
        a.gp[x + c] = (one + sm) * lpd + a.lpo[x + c];
        a.gpo[x + c] = -sm * lpd;
    }
    a.gmu[c] += lms;
}

}  // namespace

}  // namespace ARW_NS

using namespace ARW_NS;

extern "C" {

int32_t ARW_API(acoustic_substep_adjoint)(const Dom* d, const arw_params* prm,
                                          double dtau, int32_t first,
                                          const ARW_API(acoustic_adj)* adj,
                                          real* work, void* stream)
{
    int32_t rc = lib_check_domain(d);
    if (rc) return rc;
    if (!prm || !adj || !work) return lib_fail(ARW_E_ARG, "null argument");
    if ((rc = lib_check_prec(prm, d))) return rc;
    const real* lin[] = {adj->l_u_pp, adj->l_v_pp, adj->l_w_pp, adj->l_th_pp,
                         adj->l_ph_pp, adj->l_mu_pp, adj->l_p_pp,
                         adj->l_p_pp_old, adj->l_al_pp};
    real* gout[] = {adj->g_u_pp, adj->g_v_pp, adj->g_w_pp, adj->g_th_pp,
                    adj->g_ph_pp, adj->g_mu_pp, adj->g_p_pp, adj->g_p_pp_old,
                    adj->g_al_pp, adj->g_ru_t, adj->g_rv_t, adj->g_rw_t,
                    adj->g_rth_t, adj->g_rph_t, adj->g_rmu_t};
    for (const real* p : lin)
        if (!p) return lib_fail(ARW_E_ARG, "null output adjoint");
    for (real* p : gout)
        if (!p) return lib_fail(ARW_E_ARG, "null input adjoint");
    hipStream_t st = (hipStream_t)stream;
    const int nz = d->nz, ny = d->ny, nx = d->nx;
    const size_t F = (size_t)(nz + 1) * (ny + 1) * (nx + 1);
    Fw fw;
    for (int w = 0; w < W_COUNT; ++w) fw.s[w] = work + (size_t)w * F;
    // forward recompute on copies of X (arw_acoustic_substep's sequence)
    Dom dc = *d;
    real* src[] = {d->u_pp, d->v_pp, d->w_pp, d->th_pp, d->ph_pp, d->mu_pp,
                   d->p_pp, d->p_pp_old, d->al_pp};
    const size_t n3 = (size_t)nz * ny * nx, nw = (size_t)(nz + 1) * ny * nx;
    const size_t sizes[] = {(size_t)nz * ny * (nx + 1), (size_t)nz * (ny + 1) * nx,
                            nw, n3, nw, (size_t)ny * nx, n3, n3, n3};
    for (int q = 0; q < 9; ++q)
        if (hipMemcpyAsync(fw.s[q], src[q], sizes[q] * RB,
                           hipMemcpyDeviceToDevice, st) != hipSuccess)
            return lib_fail(ARW_E_LAUNCH, "adjoint: state copy failed");
    dc.u_pp = fw.s[W_U]; dc.v_pp = fw.s[W_V]; dc.w_pp = fw.s[W_W];
    dc.th_pp = fw.s[W_TH]; dc.ph_pp = fw.s[W_PH]; dc.mu_pp = fw.s[W_MU];
    dc.p_pp = fw.s[W_P]; dc.p_pp_old = fw.s[W_PO]; dc.al_pp = fw.s[W_AL];
    dc.ww_pp = fw.s[W_WW];
    Sc s = make_sc(prm, dtau, first);
    s.use_pold = 1;
    if ((rc = lib_launched(run_uv(dc, s, st), "k_uv (adjoint recompute)")))
        return rc;
    if (hipMemcpyAsync(dc.p_pp_old, dc.p_pp, n3 * RB, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return lib_fail(ARW_E_LAUNCH, "adjoint: p_pp_old copy failed");
    if ((rc = lib_launched(run_column(dc, s, dc.p_pp, 1, 1, st),
                           "k_column (adjoint recompute)")))
        return rc;
    fw.M1 = dc.mu_pp; fw.T1 = dc.th_pp; fw.W1 = dc.w_pp; fw.AL1 = dc.al_pp;
    Adj a;
    a.lu = adj->l_u_pp; a.lv = adj->l_v_pp; a.lw = adj->l_w_pp;
    a.lth = adj->l_th_pp; a.lph = adj->l_ph_pp; a.lmu = adj->l_mu_pp;
    a.lp = adj->l_p_pp; a.lpo = adj->l_p_pp_old; a.lal = adj->l_al_pp;
    a.lww = adj->l_ww_pp;
    a.gu = adj->g_u_pp; a.gv = adj->g_v_pp; a.gw = adj->g_w_pp;
    a.gth = adj->g_th_pp; a.gph = adj->g_ph_pp; a.gmu = adj->g_mu_pp;
    a.gp = adj->g_p_pp; a.gpo = adj->g_p_pp_old; a.gal = adj->g_al_pp;
    a.gru = adj->g_ru_t; a.grv = adj->g_rv_t; a.grw = adj->g_rw_t;
    a.grth = adj->g_rth_t; a.grph = adj->g_rph_t; a.grmu = adj->g_rmu_t;
    const int cols = ny * nx;
    // the damper and dt follow the forward launcher's Sc
    const dim3 lev((cols + 255) / 256, nz + 1), colg((cols + 63) / 64);
    hipLaunchKernelGGL(k_adj_p1, lev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_p1"))) return rc;
    hipLaunchKernelGGL(k_adj_s1, colg, dim3(64), 0, st, *d, s, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_s1"))) return rc;
    hipLaunchKernelGGL(k_adj_p2, lev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_p2"))) return rc;
    hipLaunchKernelGGL(k_adj_p3, lev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_p3"))) return rc;
    hipLaunchKernelGGL(k_adj_s2, colg, dim3(64), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_s2"))) return rc;
    const int nf = ny * (nx + 1) + (ny + 1) * nx;
    const dim3 flev((nf + 255) / 256, nz);
    hipLaunchKernelGGL(k_adj_f1, flev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_f1"))) return rc;
    hipLaunchKernelGGL(k_adj_f2, flev, dim3(256), 0, st, *d, s, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_f2"))) return rc;
    hipLaunchKernelGGL(k_adj_cells, dim3((cols + 127) / 128), dim3(128), 0, st,
                       *d, s, a, fw);
    return lib_launched(hipGetLastError(), "k_adj_cells");
}

int32_t ARW_API(acoustic_adjoint_work)(int32_t nz, int32_t ny, int32_t nx,
                                       int64_t* elements)
{
    if (!elements || nz < 3 || ny < 1 || nx < 1)
        return lib_fail(ARW_E_ARG, "bad extents");
    *elements = (int64_t)W_COUNT * (nz + 1) * (ny + 1) * (nx + 1);
    return ARW_OK;
}

}  // extern "C"
