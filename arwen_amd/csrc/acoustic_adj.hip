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
    W_LM1, W_LDX, W_LDY,                                     // lambda(mu''), ld
    W_RM, W_RPH, W_RGX, W_RGY, W_RMX, W_RMY, W_RAX, W_RAY,   // reference pieces
    W_RHS, W_PRH, W_WS,
    W_COUNT
};

struct Adj {
    const real *lu, *lv, *lw, *lth, *lph, *lmu, *lp, *lpo, *lal, *lww;
    real *gu, *gv, *gw, *gth, *gph, *gmu, *gp, *gpo, *gal;
    real *gru, *grv, *grw, *grth, *grph, *grmu;
    int acc;                 // tendency adjoints add to their outputs
};

// a tendency adjoint: written, or added to what the output holds
__device__ __forceinline__ void put_t(const Adj& a, real* g, int x, real v)
{
    g[x] = (a.acc ? g[x] : R(0.0)) + v;
}

struct Fw {                 // the substep's outputs, and scratch
    const real *M1, *T1, *W1, *AL1, *U1, *V1, *WW1;
    real* s[W_COUNT];
};

__device__ __forceinline__ real th_ref(const Dom& d, int k, int c)
{
    return (d.base3d ? d.thb[k * d.ny * d.nx + c] : d.thb[k]) + d.thp[k * d.ny * d.nx + c];
}

__device__ __forceinline__ real mut_of(const Dom& d, int c)
{
    return d.mub2d[c] + d.mup[c];
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

// The implicit-w system's Thomas factors of a stage (they depend on the
// stage state and dtau only): al = 1/(b - a gprev) and a of rows 1..nz,
// and gprev = c al of rows 1..nz-1 -- the same expressions k_adj_s1
// evaluates, so a sweep reading them is bit-identical to one forming them.
enum { LU_AL, LU_A, LU_GP, LU_COUNT };

__global__ __launch_bounds__(64) void k_adj_lu(Dom d, Sc s, real* lu)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ny * d.nx) return;
    const ColCtx C(d, s, c, R(0.0));
    const int st = C.st, nz = d.nz;
    const size_t F = (size_t)(nz + 1) * st;
    const real one = R(1.0);
    real gprev = R(0.0);
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
        gprev = cc * al;
        lu[LU_AL * F + k * st + c] = al;
        lu[LU_A * F + k * st + c] = ak;
        lu[LU_GP * F + k * st + c] = gprev;
    }
    const real bn = one + R(2.0) * C.cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                    * c2a_k / (C.chm(nz - 1) * C.cfm(nz));
    const real an = C.a(nz);
    lu[LU_AL * F + nz * st + c] = one / (bn - an * gprev);
    lu[LU_A * F + nz * st + c] = an;
}

// S1 with the stage's factors: the two sweeps only.
__global__ __launch_bounds__(64) void k_adj_s1_lu(Dom d, Fw f, const real* lu)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int st = d.ny * d.nx, nz = d.nz;
    if (c >= st) return;
    const size_t F = (size_t)(nz + 1) * st;
    const real* __restrict__ LWIN = f.s[W_LPH1];
    real* __restrict__ LW = f.s[W_LW];
    real vprev = R(0.0);
    for (int k = 1; k <= nz; ++k) {
        const real gp = k >= 2 ? lu[LU_GP * F + (k - 1) * st + c] : R(0.0);
        const real v = LWIN[k * st + c] - gp * vprev;
        vprev = v;
        LW[k * st + c] = v;
    }
    real carry = R(0.0);
    for (int k = nz; k >= 1; --k) {
        const int x = k * st + c;
        const real v = lu[LU_AL * F + x] * (LW[x] + carry);
        LW[x] = v;
        carry = -lu[LU_A * F + x] * v;
    }
    f.s[W_L0][c] = LWIN[c] + carry;
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
    put_t(a, a.grw, x, row ? dt * la_k : R(0.0));
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
    put_t(a, a.grph, x, k >= 1 ? dt * lr2(k) : R(0.0));
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
        put_t(a, a.grth, x, dt * l1);
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
// divergence adjoint.  Split over KG level groups (contiguous level
// chunks for the suffix sums, whose chunk totals meet in LDS).
__global__ __launch_bounds__(64 * KG) void k_adj_s2(Dom d, Sc s, Adj a, Fw f)
{
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const int st = d.ny * d.nx, nz = d.nz;
    const bool on = c < st;
    const real half = R(0.5), one = R(1.0), dt = s.dtau, eps = s.epssm;
    const real* __restrict__ MC1 = f.s[W_MC1];
    const real* __restrict__ MC2 = f.s[W_MC2];
    const real* __restrict__ MC3 = f.s[W_MC3];
    const real* __restrict__ LWW = f.s[W_LWW];
    real* __restrict__ LDVP = f.s[W_LDVP];
    real m[2] = {R(0.0), R(0.0)};                   // sum MC1 + MC3, sum MC2
    if (on)
        for (int k = ty; k <= nz; k += KG) {
            m[0] += MC1[k * st + c];
            m[1] += MC2[k * st + c];
            if (k < nz) m[0] += MC3[k * st + c];
        }
    group_sums(m);
    const real lM1 = (on ? a.lmu[c] : R(0.0)) + m[0]
                     + half * (one + eps) * m[1];
    const real lM0old = half * (one - eps) * m[1];
    // linc[k] = -sum_{m = k+1}^{nz-1} LWW[m]: chunk [k0, k1) of this group,
    // offset by the totals of the chunks above it
    const int ch = (nz + KG - 1) / KG;
    const int k0 = ty * ch < nz ? ty * ch : nz;
    const int k1 = k0 + ch < nz ? k0 + ch : nz;
    real tot[KG] = {};
    real own = R(0.0);
    if (on)
        for (int k = k0; k < k1; ++k) own += LWW[k * st + c];
    {
        __shared__ real sh[KG][64];
        sh[ty][threadIdx.x] = own;
        __syncthreads();
        for (int g = 0; g < KG; ++g) tot[g] = sh[g][threadIdx.x];
        __syncthreads();
    }
    real run = R(0.0);
    for (int g = KG - 1; g > ty; --g) run += tot[g];
    real p[1] = {R(0.0)};                           // sum dnw c1h linc
    if (on)
        for (int k = k1 - 1; k >= k0; --k) {
            const real linc = k < nz - 1 ? -run : R(0.0);
            LDVP[k * st + c] = d.dnw[k] * linc;
            p[0] += d.dnw[k] * d.c1h[k] * linc;
            run += LWW[k * st + c];
        }
    group_sums(p);
    const real ldmdt = p[0] + dt * lM1;
    if (!on) return;
    for (int k = k0; k < k1; ++k)
        LDVP[k * st + c] = LDVP[k * st + c] + d.dnw[k] * ldmdt;
    if (ty != 0) return;
    f.s[W_LM1][c] = lM1;
    put_t(a, a.grmu, c, ldmdt);
    a.gmu[c] = lM1 + lM0old;
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
    put_t(a, isu ? a.gru : a.grv, fg.out0 + k * fg.ostride, dt * l1);
    if (!fg.own) return;
    const real ldup = has_dup ? lam1(dg) : R(0.0);
    const real ld = -dt * l1 - dt * ldup;
    f.s[isu ? W_LDX : W_LDY][k * st + fg.slot] = ld;
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

// Cell pass: gather the face pieces of the advance_uv transpose.  Split
// over KG level groups; phi'' (w levels) gathers the bracket terms of the
// two mass levels it bounds.
__global__ __launch_bounds__(64 * KG) void k_adj_cells(Dom d, Sc s, Adj a, Fw f)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const bool on = c < st;
    const int cc = on ? c : 0;
    const int j = cc / nx, i = cc - j * nx;
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
    auto lb = [&](int k) {
        const int x = k * st;
        return (LBX[x + cc] - LBX[x + ip]) + (LBY[x + cc] - LBY[x + jp]);
    };
    real lms[1] = {R(0.0)};
    if (on)
        for (int k = ty; k <= nz; k += KG) {
            const int x = k * st;
            real gph = R(0.0);
            if (k < nz) {
                // the face mass sums: -1/2 c1h t of the faces the cell touches
                lms[0] += -R(0.5) * d.c1h[k] * ((LTX[x + c] + LTX[x + ip])
                                                + (LTY[x + c] + LTY[x + jp]));
                auto alt = [&](int q) { return d.alt[x + q]; };
                // face c is the cell's own (A side), face ip/jp has it as B
                real lpd = (alt(c) + alt(im)) * LBX[x + c]
                           - (alt(ip) + alt(c)) * LBX[x + ip]
                           + (alt(c) + alt(jm)) * LBY[x + c]
                           - (alt(jp) + alt(c)) * LBY[x + jp];
                lpd += LSX[x + c] + LSX[x + ip] + LSY[x + c] + LSY[x + jp];
                gph += lb(k);
                real gal = R(0.0);
                if (d.base3d) {
                    auto pb = [&](int q) { return d.pb[x + q]; };
                    gal = (pb(c) - pb(im)) * LBX[x + c] + (pb(ip) - pb(c)) * LBX[x + ip]
                          + (pb(c) - pb(jm)) * LBY[x + c] + (pb(jp) - pb(c)) * LBY[x + jp];
                }
                a.gal[x + c] = gal;
                a.gp[x + c] = (one + sm) * lpd + a.lpo[x + c];
                a.gpo[x + c] = -sm * lpd;
            }
            if (k >= 1) gph += lb(k - 1);
            a.gph[x + c] += gph;
        }
    group_sums(lms);
    if (on && ty == 0) a.gmu[c] += lms[0];
}

// ---------------------------------------------------------------------------
// The substep's dependence on the stage reference state
// ---------------------------------------------------------------------------
//
// The stage fields enter the substep as coefficients: c2a = gamma p /
// alpha_t and alpha_t of the stage EOS, the column mass mu_t = mub + mu'
// (C(mu_t) of the implicit w system, of the phi'' update and of the
// linearised EOS, the face masses of advance_uv), theta = thb + theta' (the
// theta'' fluxes, the EOS), the full phi (the Omega'' term of the phi''
// right-hand side, the PGF's phi' difference, the damper's heights), w (the
// damper's reference) and u, v (dmdt_rf).  The kernels below add J^T lam
// with respect to those to per-stage accumulators, which the caller clears
// at the stage start; arw_acoustic_reference_gradients then maps c2a and
// alpha_t through the stage EOS once per stage.  The same restated forward
// terms as the mass-column kernels above, one thread per point it writes.

struct RefAcc {
    real *mut, *c2a, *alt, *thr, *ph, *w, *grf;
    int clear;               // this call writes the accumulators
};

// the stage-field gradients the accumulators map to
struct RefOut {
    real *gu, *gv, *gw, *gthp, *gphp, *gmup;
    real *gqv;               // moist: the stage qv's gradient
    int acc;                 // add to them instead of writing
};

__device__ __forceinline__ void put_o(const RefOut& o, real* g, int x, real v)
{
    g[x] = (o.acc ? g[x] : R(0.0)) + v;
}

__device__ __forceinline__ void put_r(const RefAcc& r, real* g, int x, real v)
{
    g[x] = (r.clear ? R(0.0) : g[x]) + v;
}

// the damper at w level k: inside the layer?, dw, its sine argument
__device__ __forceinline__ bool damp_at(const Dom& d, const Sc& s, int k,
                                        int c, real& dw, real& arg, real& sn)
{
    const real g = G_F;
    const real hbot = phi_full(d, d.nz, c) / g - s.zdamp;
    const real hk = phi_full(d, k, c) / g;
    arg = R(1.5707963) * (hk - hbot) / s.zdamp;
    sn = (real)sin((double)arg);
    dw = s.dampmag * sn * sn;
    return s.damp && k >= 1 && (hk >= hbot);
}

// (column, w level): r_ (the phi'' rhs before its C(mu_t) division), the
// predicted phi'' the w accumulation differences, and the implicit solve's
// output before the damper -- each read by k_ref_lev at up to four levels
__global__ __launch_bounds__(256) void k_ref_pre(Dom d, Sc s, Fw f)
{
    int c, k;
    if (!col_level(d, c, k)) return;
    const ColCtx C(d, s, c, f.M1[c]);
    const int st = C.st, nz = d.nz, x = k * st + c;
    const real half = R(0.5), one = R(1.0), dt = s.dtau, g = G_F, eps = s.epssm;
    const real* WW1 = f.WW1;
    auto wd = [&](int kk) {                              // kk in 0..nz-1
        return half * (WW1[(kk + 1) * st + c] + WW1[kk * st + c]) * d.rdnw[kk]
               * (phi_full(d, kk + 1, c) - phi_full(d, kk, c));
    };
    real r = R(0.0), rhs = R(0.0);
    if (k >= 1) {
        r = dt * (d.rph_t[x] + half * g * (one - eps) * d.w_pp[x]);
        if (k < nz) r = r - dt * (d.fnm[k] * wd(k) + d.fnp[k] * wd(k - 1));
        if (!(k == nz && s.top_lid)) rhs = d.ph_pp[x] + r / C.cfm(k);
    }
    f.s[W_RHS][x] = r;
    f.s[W_PRH][x] = rhs;
    real ws = f.W1[x];
    real dw, arg, sn;
    if (damp_at(d, s, k, c, dw, arg, sn))
        ws = ws * (one + dw) + dw * (C.cfm(k) * d.w[x]);
    f.s[W_WS][x] = ws;
}

// (column, level 0..nz): the EOS, phi'' update, damper, Thomas solve,
// w accumulation, rhs, theta'' flux and t2 terms.  Level pieces of the
// column-mass gradient go to RM, the damper's top-height pieces to RPH.
__global__ __launch_bounds__(256) void k_ref_lev(Dom d, Sc s, Adj a, Fw f,
                                                 RefAcc r)
{
    int c, k;
    if (!col_level(d, c, k)) return;
    const ColCtx C(d, s, c, f.M1[c]);
    const int st = C.st, nz = d.nz, ny = d.ny, nx = d.nx, x = k * st + c;
    const real half = R(0.5), one = R(1.0), dt = s.dtau, g = G_F,
               eps = s.epssm, cof = C.cof;
    const real* LW = f.s[W_LW];
    const real* LRHS = f.s[W_LRHS];
    const real* LT2 = f.s[W_LT2];
    const real* GGF = f.s[W_GGF];
    const real* WW1 = f.WW1;
    const real* RR = f.s[W_RHS];
    const real* PRH = f.s[W_PRH];
    const real* WS = f.s[W_WS];
    auto mu = [&](int kk) { return LW[kk * st + c]; };   // lambda, rows 1..nz
    auto wsol = [&](int kk) { return WS[kk * st + c]; }; // pre-damper solve
    auto rhs = [&](int kk) -> real {                     // predicted phi''
        return (kk < 1 || kk > nz) ? R(0.0) : PRH[kk * st + c];
    };
    auto P0 = [&](int kk) { return d.ph_pp[kk * st + c]; };
    auto dup = [&](int rw) {
        return (one + eps) * (rhs(rw + 1) - rhs(rw))
               + (one - eps) * (P0(rw + 1) - P0(rw));
    };
    auto ddn = [&](int rw) {
        return (one + eps) * (rhs(rw) - rhs(rw - 1))
               + (one - eps) * (P0(rw) - P0(rw - 1));
    };
    auto lr2 = [&](int kk) -> real {
        if (kk < 1 || kk > nz || (kk == nz && s.top_lid)) return R(0.0);
        return LRHS[kk * st + c] / C.cfm(kk);
    };
    auto lwd = [&](int kk) -> real {
        real v = R(0.0);
        if (kk >= 1) v += -dt * d.fnm[kk] * lr2(kk);
        if (kk + 1 <= nz - 1) v += -dt * d.fnp[kk + 1] * lr2(kk + 1);
        return v;
    };
    auto gg = [&](int kk, int cc) { return GGF[kk * st + cc]; };

    real gc2a = R(0.0), galt = R(0.0), gthr = R(0.0), gchm = R(0.0);
    real gcfm = R(0.0), gph = R(0.0), gw = R(0.0), mpc = R(0.0), top = R(0.0);
    const int r1 = k + 1;
    const bool arow1 = (r1 >= 2 && r1 <= nz - 1) || (r1 == nz && !s.top_lid);
    if (k < nz) {
        const real thr = th_ref(d, k, c), alt = d.alt[x], c2a = C.c2a(k);
        const real chm = C.chm(k), chmp = C.chmp(k);
        // linearised EOS
        const real q = alt / (chmp * thr);
        const real X = f.T1[x] - d.c1h[k] * C.M1 * thr;
        const real lPP = a.lp[x];
        const real lAL = a.lal[x] - c2a * lPP;
        gc2a += lPP * (q * X - f.AL1[x]);
        const real gq = lPP * c2a * X;
        galt += gq / (chmp * thr) - lAL * d.c1h[k] * C.M1 / chmp;
        const real gchmp = -gq * q / chmp - lAL * f.AL1[x] / chmp;
        gthr += -gq * q / thr + lPP * c2a * q * (-d.c1h[k] * C.M1);
        mpc += d.c1h[k] * gchmp;
        // Thomas entries that hold level k's c2a / C(mu_t)
        if (arow1) {
            const real ga = -mu(r1) * wsol(k), av = C.a(r1);
            gc2a += ga * av / c2a;
            gchm += -ga * av / chm;
        }
        if (k >= 1) {
            const real gb = -mu(k) * wsol(k);
            const real t1 = cof * d.rdn[k] * d.rdnw[k] * c2a / (chm * C.cfm(k));
            gc2a += gb * t1 / c2a;
            gchm += -gb * t1 / chm;
            const real gcv = -mu(k) * wsol(k + 1);
            const real cv = -cof * d.rdn[k] * d.rdnw[k] * c2a / (chm * C.cfm(k + 1));
            gc2a += gcv * cv / c2a;
            gchm += -gcv * cv / chm;
        }
        if (r1 <= nz - 1) {
            const real gb = -mu(r1) * wsol(r1);
            const real t2b = cof * d.rdn[r1] * d.rdnw[k] * c2a / (chm * C.cfm(r1));
            gc2a += gb * t2b / c2a;
            gchm += -gb * t2b / chm;
        }
        if (k == nz - 1) {
            const real gb = -mu(nz) * wsol(nz);
            const real t1n = R(2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1] * c2a
                             / (chm * C.cfm(nz));
            gc2a += gb * t1n / c2a;
            gchm += -gb * t1n / chm;
        }
        // w accumulation: the PGF and theta'' terms of rows k and k+1
        const real tv = half * ((one + eps) * f.T1[x] + (one - eps) * d.th_pp[x])
                        / ((d.c1h[k] * (C.mut + C.M1) + d.c2h[k]) * thr);
        if (k >= 1) {
            const real la = mu(k), du = dup(k);
            const real cup = half * dt * g * d.rdn[k] * c2a * d.rdnw[k] / chm;
            gc2a += la * du * cup / c2a;
            gchm += -la * du * cup / chm;
            gc2a += la * dt * g * d.rdn[k] * alt * tv;
            galt += la * dt * g * d.rdn[k] * c2a * tv;
        }
        if (r1 <= nz - 1) {
            const real la = mu(r1), dd = ddn(r1);
            const real cdn = half * dt * g * d.rdn[r1] * c2a * d.rdnw[k] / chm;
            gc2a += -la * dd * cdn / c2a;
            gchm += la * dd * cdn / chm;
            gc2a -= la * dt * g * d.rdn[r1] * alt * tv;
            galt -= la * dt * g * d.rdn[r1] * c2a * tv;
        }
        if (k == nz - 1 && !s.top_lid) {
            const real la = mu(nz);
            const real cb = dt * g * c2a * d.rdnw[nz - 1] * d.rdnw[nz - 1] / chm;
            const real bt = cb * ddn(nz);
            gc2a += -la * bt / c2a;
            gchm += la * bt / chm;
            gc2a += -la * dt * g * R(2.0) * d.rdnw[nz - 1] * alt * tv;
            galt += -la * dt * g * R(2.0) * d.rdnw[nz - 1] * c2a * tv;
        }
        // t2 = theta''-average / (C(mu_t + mu'') theta)
        const real chs = d.c1h[k] * (C.mut + C.M1) + d.c2h[k];
        mpc += -tv * d.c1h[k] / chs * LT2[x];
        gthr += -tv / thr * LT2[x];
        // theta of the theta'' fluxes: vertical (Omega'') and horizontal
        if (k >= 1)
            gthr += (d.rdnw[k - 1] * gg(k - 1, c) - d.rdnw[k] * gg(k, c))
                    * WW1[x] * d.fnm[k];
        if (r1 <= nz - 1)
            gthr += (d.rdnw[k] * gg(k, c) - d.rdnw[r1] * gg(r1, c))
                    * WW1[r1 * st + c] * d.fnp[r1];
        const int j = c / nx, i = c - j * nx;
        const int cim = j * nx + (i == 0 ? nx - 1 : i - 1);
        const int cip = j * nx + (i + 1 == nx ? 0 : i + 1);
        const int cjm = (j == 0 ? ny - 1 : j - 1) * nx + i;
        const int cjp = (j + 1 == ny ? 0 : j + 1) * nx + i;
        const real* U1 = f.U1;
        const real* V1 = f.V1;
        const real g0 = gg(k, c);
        gthr += half * s.rdx * (U1[iU(k, j, i + 1, ny, nx)] - U1[iU(k, j, i, ny, nx)]) * g0
                + half * s.rdy * (V1[iV(k, j + 1, i, ny, nx)] - V1[iV(k, j, i, ny, nx)]) * g0;
        gthr += half * s.rdx * U1[iU(k, j, i == 0 ? nx : i, ny, nx)] * gg(k, cim)
                - half * s.rdx * U1[iU(k, j, i + 1 == nx ? 0 : i + 1, ny, nx)] * gg(k, cip)
                + half * s.rdy * V1[iV(k, j == 0 ? ny : j, i, ny, nx)] * gg(k, cjm)
                - half * s.rdy * V1[iV(k, j + 1 == ny ? 0 : j + 1, i, ny, nx)] * gg(k, cjp);
        put_r(r, r.c2a, x, gc2a);
        put_r(r, r.alt, x, galt);
        put_r(r, r.thr, x, gthr);
        mpc += d.c1h[k] * gchm;
    }
    // w level k: C(mu_t) of the Thomas rows, the rhs, the phi'' update and
    // the damper; the Omega'' term's phi difference
    const real cfmk = C.cfm(k);
    if (arow1) {
        const real ga = -mu(r1) * wsol(k);
        gcfm += -ga * C.a(r1) / cfmk;
    }
    if (k >= 1 && k <= nz - 1) {
        const real gb = -mu(k) * wsol(k);
        const real t1 = cof * d.rdn[k] * d.rdnw[k] * C.c2a(k) / (C.chm(k) * cfmk);
        const real t2b = cof * d.rdn[k] * d.rdnw[k - 1] * C.c2a(k - 1)
                         / (C.chm(k - 1) * cfmk);
        gcfm += -gb * (t1 + t2b) / cfmk;
    }
    if (k >= 2) {
        const int rm = k - 1;
        const real gcv = -mu(rm) * wsol(k);
        const real cv = -cof * d.rdn[rm] * d.rdnw[rm] * C.c2a(rm) / (C.chm(rm) * cfmk);
        gcfm += -gcv * cv / cfmk;
    }
    if (k == nz) {
        const real gb = -mu(nz) * wsol(nz);
        const real t1n = R(2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1] * C.c2a(nz - 1)
                         / (C.chm(nz - 1) * cfmk);
        gcfm += -gb * t1n / cfmk;
    }
    if (k >= 1) {
        const real lr = (k == nz && s.top_lid) ? R(0.0) : LRHS[x];
        gcfm += -lr * RR[x] / (cfmk * cfmk);
        real ph1 = a.lph[x];
        if (k < nz) ph1 += d.rdnw[k] / C.chmp(k) * C.lAL(a, k);
        ph1 += -d.rdnw[k - 1] / C.chmp(k - 1) * C.lAL(a, k - 1);
        const real cfs = C.cfs(k);
        mpc += -C.K * f.W1[x] * d.c1f[k] / (cfs * cfs) * ph1;
        real dw, arg, sn;
        if (damp_at(d, s, k, c, dw, arg, sn)) {
            const real lk = a.lw[x] + C.K / cfs * ph1;
            const real wk = d.w[x], cw = cfmk * wk;
            const real ws = wsol(k);
            gw += -lk * dw * cfmk / (one + dw);
            gcfm += -lk * dw * wk / (one + dw);
            const real gdw = -lk * (cw + ws) / ((one + dw) * (one + dw));
            const real garg = gdw * s.dampmag * R(2.0) * sn * (real)cos((double)arg);
            const real ghk = garg * R(1.5707963) / s.zdamp;
            gph += ghk / g;
            top = -ghk / g;
        }
    }
    auto gdp = [&](int kk) {
        return lwd(kk) * half * (WW1[(kk + 1) * st + c] + WW1[kk * st + c]) * d.rdnw[kk];
    };
    if (k >= 1) gph += gdp(k - 1);
    if (k <= nz - 1) gph -= gdp(k);
    mpc += d.c1f[k] * gcfm;
    put_r(r, r.ph, x, gph);
    put_r(r, r.w, x, gw);
    f.s[W_RM][x] = mpc;
    f.s[W_RPH][x] = top;
}

// (own face, level): the advance_uv pieces -- the column-mass weight
// (RM*), the alpha_t term (RA*) and the phi' difference (RG*)
__global__ __launch_bounds__(256) void k_ref_face(Dom d, Sc s, Fw f)
{
    FaceGeo fg, dg;
    bool has_dup, isu;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (!face_geo(d, t, fg, dg, has_dup, isu) || !fg.own) return;
    const int st = d.ny * d.nx, nz = d.nz;
    const real half = R(0.5);
    const real rd = isu ? s.rdx : s.rdy;
    const real c1 = d.c1h[k];
    const int A = fg.cA, B = fg.cB, xs = k * st + fg.slot;
    const real ld = f.s[isu ? W_LDX : W_LDY][xs];
    auto pd = [&](int kk, int cc) {
        const int q = kk * st + cc;
        return d.p_pp[q] + s.smdiv * (d.p_pp[q] - d.p_pp_old[q]);
    };
    const int h = k * st, hp = (k + 1) * st;
    const real dph = (d.ph_pp[hp + A] - d.ph_pp[hp + B])
                     + (d.ph_pp[h + A] - d.ph_pp[h + B]);
    const real dpe = pd(k, A) - pd(k, B);
    real brk = dph + (d.alt[h + A] + d.alt[h + B]) * dpe;
    if (d.base3d)
        brk = brk + (d.al_pp[h + A] + d.al_pp[h + B]) * (d.pb[h + A] - d.pb[h + B]);
    f.s[isu ? W_RMX : W_RMY][xs] = half * rd * c1 * brk * ld;
    f.s[isu ? W_RAX : W_RAY][xs] = f.s[isu ? W_LBX : W_LBY][xs] * dpe;
    auto S = [&](int kk) { return pd(kk, A) + pd(kk, B); };
    auto dpn = [&](int kf) -> real {
        if (kf == 0) return half * (d.cf1 * S(0) + d.cf2 * S(1) + d.cf3 * S(2));
        if (kf < nz) return half * (d.fnm[kf] * S(kf) + d.fnp[kf] * S(kf - 1));
        if (s.top_lid)
            return half * (d.cf1 * S(nz - 1) + d.cf2 * S(nz - 2) + d.cf3 * S(nz - 3));
        return R(0.0);
    };
    f.s[isu ? W_RGX : W_RGY][xs] =
        rd * (d.rdnw[k] * (dpn(k + 1) - dpn(k)) - half * c1 * (d.mu_pp[B] + d.mu_pp[A]))
        * ld;
}

// (column, w level): gather the own-face pieces onto the cells; the
// column-mass pieces join RM
__global__ __launch_bounds__(256) void k_ref_cells(Dom d, Fw f, RefAcc r)
{
    int c, k;
    if (!col_level(d, c, k)) return;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx, x = k * st + c;
    const int j = c / nx, i = c - j * nx;
    const int ip = j * nx + (i + 1 == nx ? 0 : i + 1);
    const int jp = (j + 1 == ny ? 0 : j + 1) * nx + i;
    const real half = R(0.5);
    const real *RGX = f.s[W_RGX], *RGY = f.s[W_RGY];
    auto gphs = [&](int kk) {
        const int y = kk * st;
        return half * (RGX[y + c] - RGX[y + ip]) + half * (RGY[y + c] - RGY[y + jp]);
    };
    real gph = R(0.0);
    if (k < nz) {
        const real *RAX = f.s[W_RAX], *RAY = f.s[W_RAY];
        const real *RMX = f.s[W_RMX], *RMY = f.s[W_RMY];
        r.alt[x] += (RAX[x] + RAX[k * st + ip]) + (RAY[x] + RAY[k * st + jp]);
        f.s[W_RM][x] += half * (RMX[x] + RMX[k * st + ip])
                        + half * (RMY[x] + RMY[k * st + jp]);
        gph += gphs(k);
    }
    if (k >= 1) gph += gphs(k - 1);
    r.ph[x] += gph;
}

// per column: the level pieces of the column-mass gradient and the
// damper's top height; this substep's dmdt_rf adjoint joins grf (the stage
// reference u, v and mass it weights are mapped once per stage)
__global__ __launch_bounds__(64 * KG) void k_ref_col(Dom d, Sc s, Fw f, RefAcc r)
{
    const int nz = d.nz, st = d.ny * d.nx;
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const bool on = c < st;
    real v[2] = {R(0.0), R(0.0)};                   // mass, top height
    if (on)
        for (int k = ty; k <= nz; k += KG) {
            v[0] += f.s[W_RM][k * st + c];
            v[1] += f.s[W_RPH][k * st + c];
        }
    group_sums(v);
    if (!on || ty != 0) return;
    r.ph[nz * st + c] += v[1];
    put_r(r, r.mut, c, v[0]);
    put_r(r, r.grf, c, s.dtau * f.s[W_LM1][c]);
}

// ---- c2a and alpha_t through the stage EOS (dry, hypsometric_opt = 1) ----
enum { G_DPH, G_MU3, G_COUNT };

__global__ void k_refg_eos(Dom d, RefAcc r, RefOut o, real* w)
{
    const int nz = d.nz, st = d.ny * d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nz * st) return;
    const int k = t / st, c = t - k * st;
    const size_t F = (size_t)(nz + 1) * (d.ny + 1) * (d.nx + 1);
    const real p = d.p[t], alt = d.alt[t];
    const real c2a = (GAMMA_F * p) / alt;
    const real gp = r.c2a[t] * GAMMA_F / alt;
    const real galt = r.alt[t] - r.c2a[t] * c2a / alt;
    const real th = th_ref(d, k, c);
    const real mh = d.c1h[k] * (d.mub2d[c] + d.mup[c]) + d.c2h[k];
    const real ga = galt - gp * (GAMMA_F * p / alt);
    put_o(o, o.gthp, t, r.thr[t] + gp * (GAMMA_F * p / th));
    if (d.qv) {            // theta_m = theta (1 + Rv/Rd qv) of the moist EOS
        const real thm = th * (R(1.0) + RVOVRD_F * d.qv[t]);
        put_o(o, o.gqv, t, gp * (GAMMA_F * p / thm) * th * RVOVRD_F);
    }
    w[G_DPH * F + t] = ga * (-d.rdnw[k] / mh);
    w[G_MU3 * F + t] = ga * (-alt * d.c1h[k] / mh);
}

__global__ void k_refg_php(Dom d, RefAcc r, RefOut o, const real* w)
{
    const int nz = d.nz, st = d.ny * d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (nz + 1) * st) return;
    const int k = t / st;
    real g = r.ph[t];
    if (k >= 1) g += w[t - st];
    if (k <= nz - 1) g -= w[t];
    put_o(o, o.gphp, t, g);
    put_o(o, o.gw, t, r.w[t]);
}

// the stage's dmdt_rf = sum_k dnw (divergence of the coupled u, v): its
// share for the substep-summed grf, on the faces (u, v) ...
__global__ __launch_bounds__(256) void k_refg_dmdt(Dom d, Sc s, RefAcc r,
                                                   RefOut o)
{
    FaceGeo fg, dg;
    bool has_dup, isu;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (!face_geo(d, t, fg, dg, has_dup, isu)) return;
    const real half = R(0.5);
    const real rd = isu ? s.rdx : s.rdy;
    const real c1 = d.c1h[k], c2 = d.c2h[k];
    real v = R(0.0);
    if (fg.hasE)
        v += rd * (c1 * (half * (mut_of(d, fg.cE) + mut_of(d, fg.cEp))) + c2)
             * (d.dnw[k] * r.grf[fg.cE]);
    if (fg.hasW)
        v -= rd * (c1 * (half * (mut_of(d, fg.cWm) + mut_of(d, fg.cW))) + c2)
             * (d.dnw[k] * r.grf[fg.cW]);
    put_o(o, isu ? o.gu : o.gv, fg.out0 + k * fg.ostride, v);
}

// ... and in the column mass (with the EOS's mu' column sum)
__global__ __launch_bounds__(64 * KG) void k_refg_mu(Dom d, Sc sc, RefAcc r,
                                                    RefOut o, const real* w)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const bool on = c < st;
    const int cc = on ? c : 0;
    const size_t F = (size_t)(nz + 1) * (ny + 1) * (nx + 1);
    const int j = cc / nx, i = cc - j * nx;
    const int im = i == 0 ? nx - 1 : i - 1, ip = i + 1 == nx ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j + 1 == ny ? 0 : j + 1;
    // the EOS's mu' terms, then the column sums of c1h dnw u (v) on the
    // faces i, i+1 and the neighbours' faces (i == 0 ? nx : i), ip
    const int fu[4] = {i, i + 1, i == 0 ? nx : i, ip};
    const int fv[4] = {j, j + 1, j == 0 ? ny : j, jp};
    real v[9] = {};
    if (on)
        for (int k = ty; k < nz; k += KG) {
            v[0] += w[G_MU3 * F + (size_t)k * st + c];
            const real cd = d.c1h[k];
            for (int q = 0; q < 4; ++q) {
                v[1 + q] += cd * d.u[iU(k, j, fu[q], ny, nx)] * d.dnw[k];
                v[5 + q] += cd * d.v[iV(k, fv[q], i, ny, nx)] * d.dnw[k];
            }
        }
    group_sums(v);
    if (!on || ty != 0) return;
    const real half = R(0.5);
    const real G = r.grf[c];
    real m = half * (sc.rdx * G * v[2] + sc.rdx * r.grf[j * nx + im] * v[3])
             + half * (-sc.rdx * G * v[1] - sc.rdx * r.grf[j * nx + ip] * v[4]);
    m += half * (sc.rdy * G * v[6] + sc.rdy * r.grf[jm * nx + i] * v[7])
         + half * (-sc.rdy * G * v[5] - sc.rdy * r.grf[jp * nx + i] * v[8]);
    put_o(o, o.gmup, c, (r.mut[c] + m) + v[0]);
}

// ---- the kinematic lower boundary w(sfc) = u . grad(ht) ----------------
// (u, v) gradients for lam_w0, added to gu / gv at levels 0..2 (the
// extrapolation weights cf1..cf3), one thread per face of those levels
__global__ void k_wsfc_adj(Dom d, real hdx, real hdy, const real* l0,
                           real* gu, real* gv)
{
    const int ny = d.ny, nx = d.nx;
    const int nu = ny * (nx + 1), nv = (ny + 1) * nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (t >= nu + nv) return;
    const real cf = k == 0 ? d.cf1 : (k == 1 ? d.cf2 : d.cf3);
    const real* ht = d.ht;
    if (t < nu) {
        const int j = t / (nx + 1), fc = t - j * (nx + 1);
        real v = gu[iU(k, j, fc, ny, nx)];
        if (fc >= 1) {                  // east face of cell fc-1
            const int c = j * nx + fc - 1, ce = j * nx + (fc == nx ? 0 : fc);
            v += cf * (hdx * (ht[ce] - ht[c]) * l0[c]);
        }
        if (fc <= nx - 1) {             // west face of cell fc
            const int c = j * nx + fc, cw = j * nx + (fc == 0 ? nx - 1 : fc - 1);
            v += cf * (hdx * (ht[c] - ht[cw]) * l0[c]);
        }
        gu[iU(k, j, fc, ny, nx)] = v;
    } else {
        const int q = t - nu, g = q / nx, i = q - g * nx;
        real v = gv[iV(k, g, i, ny, nx)];
        if (g >= 1) {
            const int c = (g - 1) * nx + i, cn = (g == ny ? 0 : g) * nx + i;
            v += cf * (hdy * (ht[cn] - ht[c]) * l0[c]);
        }
        if (g <= ny - 1) {
            const int c = g * nx + i, cs = (g == 0 ? ny - 1 : g - 1) * nx + i;
            v += cf * (hdy * (ht[c] - ht[cs]) * l0[c]);
        }
        gv[iV(k, g, i, ny, nx)] = v;
    }
}

}  // namespace

}  // namespace ARW_NS

using namespace ARW_NS;

// The forward substep on copies of X in the work area (arw_acoustic_
// substep's sequence): the outputs the transpose is taken about.
static int32_t recompute(const Dom* d, const Sc& s, Fw& fw, hipStream_t st)
{
    int32_t rc;
    const int nz = d->nz, ny = d->ny, nx = d->nx;
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
    if ((rc = lib_launched(run_uv(dc, s, st), "k_uv (adjoint recompute)")))
        return rc;
    if (hipMemcpyAsync(dc.p_pp_old, dc.p_pp, n3 * RB, hipMemcpyDeviceToDevice,
                       st) != hipSuccess)
        return lib_fail(ARW_E_LAUNCH, "adjoint: p_pp_old copy failed");
    if ((rc = lib_launched(run_column(dc, s, dc.p_pp, 1, 1, st),
                           "k_column (adjoint recompute)")))
        return rc;
    fw.M1 = dc.mu_pp; fw.T1 = dc.th_pp; fw.W1 = dc.w_pp; fw.AL1 = dc.al_pp;
    fw.U1 = dc.u_pp; fw.V1 = dc.v_pp; fw.WW1 = dc.ww_pp;
    return ARW_OK;
}

static int32_t substep_adjoint(const Dom* d, const arw_params* prm,
                               double dtau, int32_t first,
                               const ARW_API(acoustic_adj)* adj,
                               const RefAcc* ref,
                               const ARW_API(acoustic_fwd)* fwd,
                               const real* lu, int acc, real* work,
                               void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
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
    Sc s = make_sc(prm, dtau, first);
    s.use_pold = 1;
    if (fwd) {            // the substep's outputs, kept by the caller
        fw.U1 = fwd->u_pp; fw.V1 = fwd->v_pp; fw.W1 = fwd->w_pp;
        fw.T1 = fwd->th_pp; fw.M1 = fwd->mu_pp; fw.AL1 = fwd->al_pp;
        fw.WW1 = fwd->ww_pp;
    } else if ((rc = recompute(d, s, fw, st))) {
        return rc;
    }
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
    a.acc = acc;
    const int cols = ny * nx;
    // the damper and dt follow the forward launcher's Sc
    const dim3 lev((cols + 255) / 256, nz + 1), colg((cols + 63) / 64);
    hipLaunchKernelGGL(k_adj_p1, lev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_p1"))) return rc;
    if (lu)
        hipLaunchKernelGGL(k_adj_s1_lu, colg, dim3(64), 0, st, *d, fw, lu);
    else
        hipLaunchKernelGGL(k_adj_s1, colg, dim3(64), 0, st, *d, s, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_s1"))) return rc;
    hipLaunchKernelGGL(k_adj_p2, lev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_p2"))) return rc;
    hipLaunchKernelGGL(k_adj_p3, lev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_p3"))) return rc;
    if (ref) {       // reads the solve's lambda before the face pass reuses it
        hipLaunchKernelGGL(k_ref_pre, lev, dim3(256), 0, st, *d, s, fw);
        if ((rc = lib_launched(hipGetLastError(), "k_ref_pre"))) return rc;
        hipLaunchKernelGGL(k_ref_lev, lev, dim3(256), 0, st, *d, s, a, fw, *ref);
        if ((rc = lib_launched(hipGetLastError(), "k_ref_lev"))) return rc;
    }
    hipLaunchKernelGGL(k_adj_s2, colg, dim3(64, KG), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_s2"))) return rc;
    const int nf = ny * (nx + 1) + (ny + 1) * nx;
    const dim3 flev((nf + 255) / 256, nz);
    hipLaunchKernelGGL(k_adj_f1, flev, dim3(256), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_f1"))) return rc;
    if (ref) {
        hipLaunchKernelGGL(k_ref_face, flev, dim3(256), 0, st, *d, s, fw);
        if ((rc = lib_launched(hipGetLastError(), "k_ref_face"))) return rc;
    }
    hipLaunchKernelGGL(k_adj_f2, flev, dim3(256), 0, st, *d, s, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_f2"))) return rc;
    hipLaunchKernelGGL(k_adj_cells, colg, dim3(64, KG), 0, st, *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_cells"))) return rc;
    if (!ref) return ARW_OK;
    hipLaunchKernelGGL(k_ref_cells, lev, dim3(256), 0, st, *d, fw, *ref);
    if ((rc = lib_launched(hipGetLastError(), "k_ref_cells"))) return rc;
    hipLaunchKernelGGL(k_ref_col, dim3((cols + 63) / 64), dim3(64, KG), 0, st,
                       *d, s, fw, *ref);
    return lib_launched(hipGetLastError(), "k_ref_col");
}

static bool ref_ok(const ARW_API(acoustic_ref)* acc)
{
    return acc && acc->mut && acc->c2a && acc->alt && acc->thr && acc->ph
           && acc->w && acc->grf;
}

static RefAcc ref_of(const ARW_API(acoustic_ref)* acc)
{
    RefAcc r;
    r.mut = acc->mut; r.c2a = acc->c2a; r.alt = acc->alt; r.thr = acc->thr;
    r.ph = acc->ph; r.w = acc->w; r.grf = acc->grf;
    r.clear = 0;
    return r;
}

extern "C" {

int32_t ARW_API(acoustic_substep_adjoint)(const Dom* d, const arw_params* prm,
                                          double dtau, int32_t first,
                                          const ARW_API(acoustic_adj)* adj,
                                          real* work, void* stream)
{
    return substep_adjoint(d, prm, dtau, first, adj, nullptr, nullptr, nullptr,
                           0, work, stream);
}

int32_t ARW_API(acoustic_substep_adjoint_ref)(
    const Dom* d, const arw_params* prm, double dtau, int32_t first,
    const ARW_API(acoustic_adj)* adj, const ARW_API(acoustic_ref)* acc,
    const ARW_API(acoustic_fwd)* fwd, const real* lu, int32_t flags,
    real* work, void* stream)
{
    if (flags & ~(ARW_ADJ_CLEAR | ARW_ADJ_ACCUMULATE))
        return lib_fail(ARW_E_ARG, "unknown adjoint flags");
    if (!ref_ok(acc)) return lib_fail(ARW_E_ARG, "null reference accumulator");
    if (fwd && !(fwd->u_pp && fwd->v_pp && fwd->w_pp && fwd->th_pp
                 && fwd->mu_pp && fwd->al_pp && fwd->ww_pp))
        return lib_fail(ARW_E_ARG, "null substep output");
    RefAcc r = ref_of(acc);
    r.clear = (flags & ARW_ADJ_CLEAR) != 0;
    return substep_adjoint(d, prm, dtau, first, adj, &r, fwd, lu,
                           (flags & ARW_ADJ_ACCUMULATE) != 0, work, stream);
}

int32_t ARW_API(acoustic_adjoint_factors)(const Dom* d, const arw_params* prm,
                                          double dtau, real* lu, void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
    if (rc) return rc;
    if (!prm || !lu) return lib_fail(ARW_E_ARG, "null argument");
    if ((rc = lib_check_prec(prm, d))) return rc;
    const Sc s = make_sc(prm, dtau, 1);
    const int cols = d->ny * d->nx;
    hipLaunchKernelGGL(k_adj_lu, dim3((cols + 63) / 64), dim3(64), 0,
                       (hipStream_t)stream, *d, s, lu);
    return lib_launched(hipGetLastError(), "k_adj_lu");
}

int32_t ARW_API(acoustic_reference_gradients)(
    const Dom* d, const arw_params* prm, const ARW_API(acoustic_ref)* acc,
    real* g_u, real* g_v, real* g_w, real* g_thp, real* g_php, real* g_mup,
    real* g_qv, int32_t flags, real* work, void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
    if (rc) return rc;
    if (!prm || !ref_ok(acc) || !g_u || !g_v || !g_w || !g_thp || !g_php
        || !g_mup || !work)
        return lib_fail(ARW_E_ARG, "null argument");
    if (flags & ~ARW_ADJ_ACCUMULATE) return lib_fail(ARW_E_ARG, "unknown adjoint flags");
    if (d->qv && !g_qv)
        return lib_fail(ARW_E_ARG, "reference adjoint: a moist state needs g_qv");
    const RefAcc r = ref_of(acc);
    RefOut o;
    o.gqv = g_qv;
    o.gu = g_u; o.gv = g_v; o.gw = g_w; o.gthp = g_thp; o.gphp = g_php;
    o.gmup = g_mup; o.acc = (flags & ARW_ADJ_ACCUMULATE) != 0;
    hipStream_t st = (hipStream_t)stream;
    const int nz = d->nz, cols = d->ny * d->nx;
    const int nm = nz * cols, nw = (nz + 1) * cols;
    hipLaunchKernelGGL(k_refg_eos, dim3((nm + 255) / 256), dim3(256), 0, st,
                       *d, r, o, work);
    if ((rc = lib_launched(hipGetLastError(), "k_refg_eos"))) return rc;
    hipLaunchKernelGGL(k_refg_php, dim3((nw + 255) / 256), dim3(256), 0, st,
                       *d, r, o, work);
    if ((rc = lib_launched(hipGetLastError(), "k_refg_php"))) return rc;
    const Sc sc = make_sc(prm, prm->dt, 1);
    const int nf = d->ny * (d->nx + 1) + (d->ny + 1) * d->nx;
    hipLaunchKernelGGL(k_refg_dmdt, dim3((nf + 255) / 256, nz), dim3(256), 0,
                       st, *d, sc, r, o);
    if ((rc = lib_launched(hipGetLastError(), "k_refg_dmdt"))) return rc;
    hipLaunchKernelGGL(k_refg_mu, dim3((cols + 63) / 64), dim3(64, KG), 0, st,
                       *d, sc, r, o, work);
    return lib_launched(hipGetLastError(), "k_refg_mu");
}

int32_t ARW_API(set_w_surface_adjoint)(const Dom* d, const arw_params* prm,
                                       const real* lam_w0, real* g_u,
                                       real* g_v, void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
    if (rc) return rc;
    if (!prm || !lam_w0 || !g_u || !g_v) return lib_fail(ARW_E_ARG, "null argument");
    if (d->nz < 3) return lib_fail(ARW_E_ARG, "set_w_surface adjoint needs nz >= 3");
    const int nf = d->ny * (d->nx + 1) + (d->ny + 1) * d->nx;
    hipLaunchKernelGGL(k_wsfc_adj, dim3((nf + 255) / 256, 3), dim3(256), 0,
                       (hipStream_t)stream, *d, (real)(0.5 / prm->dx),
                       (real)(0.5 / prm->dy), lam_w0, g_u, g_v);
    return lib_launched(hipGetLastError(), "k_wsfc_adj");
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
