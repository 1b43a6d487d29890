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
// Three passes, each a plain thread-per-column march (this is not the
// throughput path):
//   k_adj_col    mass columns: the vertical reverse of the w/phi solve,
//                the EOS, the Thomas solve, the theta'' and mu''/Omega''
//                updates; leaves per-cell adjoints for the horizontal parts
//   k_adj_faces  u and v faces: adjoint of u''/v'' (gathered from the two
//                cells of each face) and the per-face pieces of the
//                advance_uv transpose
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
    W_LBX, W_LSX, W_LBY, W_LSY, W_LMX, W_LMY,                // faces -> cells
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

__device__ __forceinline__ real phi_ref(const Dom& d, int k, int c)
{
    return (d.base3d ? d.phb[k * d.ny * d.nx + c] : d.phb[k]) + d.php[k * d.ny * d.nx + c];
}

// Mass-column reverse of advance_w_phi + diagnose_p and of the vertical
// parts of advance_mu_th.
__global__ void k_adj_col(Dom d, Sc s, Adj a, Fw f)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const real half = R(0.5), one = R(1.0), g = G_F;
    const real dt = s.dtau, eps = s.epssm;
    const real mut = d.mub2d[c] + d.mup[c];
    const real M1 = f.M1[c], M0 = d.mu_pp[c];
    real* LPH1 = f.s[W_LPH1];
    real* LW = f.s[W_LW];
    real* SA = f.s[W_SA];
    real* SG = f.s[W_SG];
    real* LRHS = f.s[W_LRHS];
    real* LT2 = f.s[W_LT2];
    real* LWD = f.s[W_LWD];
    real* LWW = f.s[W_LWW];
    real* GG = f.s[W_GG];
    real* LDVP = f.s[W_LDVP];
    auto c2a_at = [&](int k) {
        return (GAMMA_F * d.p[k * st + c]) / d.alt[k * st + c];
    };
    real lM1 = a.lmu[c];

    // phi''-adjoint column, EOS reverse; GG holds lambda(theta''_new)
    for (int k = 0; k <= nz; ++k) LPH1[k * st + c] = a.lph[k * st + c];
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        const real thr = th_ref(d, k, c), alt = d.alt[x], c2a = c2a_at(k);
        const real chmp = d.c1h[k] * (mut + M1) + d.c2h[k];
        const real q = alt / (chmp * thr);
        const real X = f.T1[x] - d.c1h[k] * M1 * thr;
        const real lPP = a.lp[x];
        GG[x] = a.lth[x] + c2a * q * lPP;
        lM1 += c2a * (-q * d.c1h[k] * thr - q * X * d.c1h[k] / chmp) * lPP;
        const real lAL = a.lal[x] - c2a * lPP;
        lM1 += (-(alt * d.c1h[k]) / chmp - f.AL1[x] * d.c1h[k] / chmp) * lAL;
        LPH1[(k + 1) * st + c] += -d.rdnw[k] / chmp * lAL;
        LPH1[x] += d.rdnw[k] / chmp * lAL;
    }
    // phi'' update: Ph1 = rhs + K W1 / cfs
    const real K = half * dt * g * (one + eps);
    LRHS[c] = R(0.0);
    LW[c] = a.lw[c];
    for (int k = 1; k <= nz; ++k) {
        const int x = k * st + c;
        const real cfs = d.c1f[k] * (mut + M1) + d.c2f[k];
        LRHS[x] = LPH1[x];
        LW[x] = a.lw[x] + K / cfs * LPH1[x];
        lM1 += -K * f.W1[x] * d.c1f[k] / (cfs * cfs) * LPH1[x];
    }
    real gph0 = LPH1[c];                       // phi''(0) passes through
    // damper
    if (s.damp) {
        const real htop = phi_ref(d, nz, c) / g;
        const real hbot = htop - s.zdamp;
        for (int k = 1; k <= nz; ++k) {
            const real hk = phi_ref(d, k, c) / g;
            if (hk >= hbot) {
                const real arg = R(1.5707963) * (hk - hbot) / s.zdamp;
                const real sn = (real)sin((double)arg);
                const real dw = s.dampmag * sn * sn;
                LW[k * st + c] = LW[k * st + c] / (one + dw);
            }
        }
    }
    // Thomas factors of the implicit w system (calc_coef_w)
    real cof = half * dt * g * (one + eps);
    cof = cof * cof;
    auto chm_at = [&](int k) { return d.c1h[k] * mut + d.c2h[k]; };
    auto cfm_at = [&](int k) { return d.c1f[k] * mut + d.c2f[k]; };
    auto a_at = [&](int k) -> real {
        if (k >= 2 && k < nz)
            return -cof * d.rdn[k] * d.rdnw[k - 1] * c2a_at(k - 1)
                   / (chm_at(k - 1) * cfm_at(k - 1));
        if (k == nz && !s.top_lid)
            return R(-2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                   * c2a_at(nz - 1) / (chm_at(nz - 1) * cfm_at(nz - 1));
        return R(0.0);
    };
    {
        real gprev = R(0.0);
        for (int k = 1; k < nz; ++k) {
            const real b = one + cof * d.rdn[k] * (
                d.rdnw[k] * c2a_at(k) / (chm_at(k) * cfm_at(k))
                + d.rdnw[k - 1] * c2a_at(k - 1) / (chm_at(k - 1) * cfm_at(k)));
            const real cc = -cof * d.rdn[k] * d.rdnw[k] * c2a_at(k)
                            / (chm_at(k) * cfm_at(k + 1));
            const real al = one / (b - a_at(k) * gprev);
            SA[k * st + c] = al;
            gprev = cc * al;
            SG[k * st + c] = gprev;
        }
        const real bn = one + R(2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                        * c2a_at(nz - 1) / (chm_at(nz - 1) * cfm_at(nz));
        SA[nz * st + c] = one / (bn - a_at(nz) * gprev);
    }
    for (int k = 1; k < nz; ++k)
        LW[(k + 1) * st + c] += -SG[k * st + c] * LW[k * st + c];
    for (int k = nz; k >= 1; --k) {
        const real v = SA[k * st + c] * LW[k * st + c];
        LW[k * st + c] = v;
        LW[(k - 1) * st + c] += -a_at(k) * v;
    }
    // w accumulation reverse
    for (int k = 0; k < nz; ++k) LT2[k * st + c] = R(0.0);
    real lmuave = R(0.0);
    a.gw[c] = R(0.0);
    for (int k = 1; k <= nz; ++k) a.gw[k * st + c] = R(0.0);
    for (int k = 0; k <= nz; ++k) a.gph[k * st + c] = R(0.0);
    for (int k = 1; k < nz; ++k) {
        const int x = k * st + c;
        const real la = LW[x];
        a.gw[x] += la;
        a.grw[x] = dt * la;
        const real cup = half * dt * g * d.rdn[k] * c2a_at(k) * d.rdnw[k] / chm_at(k);
        const real cdn = half * dt * g * d.rdn[k] * c2a_at(k - 1) * d.rdnw[k - 1]
                         / chm_at(k - 1);
        const real lup = cup * la, ldn = -cdn * la;
        LRHS[(k + 1) * st + c] += (one + eps) * lup;
        LRHS[x] -= (one + eps) * lup;
        a.gph[(k + 1) * st + c] += (one - eps) * lup;
        a.gph[x] -= (one - eps) * lup;
        LRHS[x] += (one + eps) * ldn;
        LRHS[(k - 1) * st + c] -= (one + eps) * ldn;
        a.gph[x] += (one - eps) * ldn;
        a.gph[(k - 1) * st + c] -= (one - eps) * ldn;
        LT2[x] += dt * g * d.rdn[k] * c2a_at(k) * d.alt[x] * la;
        LT2[(k - 1) * st + c] -= dt * g * d.rdn[k] * c2a_at(k - 1)
                                 * d.alt[(k - 1) * st + c] * la;
        lmuave += -dt * g * d.c1f[k] * la;
    }
    a.grw[c] = R(0.0);
    if (s.top_lid) {
        a.grw[nz * st + c] = R(0.0);
    } else {
        const int x = nz * st + c;
        const real la = LW[x];
        a.gw[x] += la;
        a.grw[x] = dt * la;
        const real cb = dt * g * c2a_at(nz - 1) * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                        / chm_at(nz - 1);
        const real ldn = -cb * la;
        LRHS[x] += (one + eps) * ldn;
        LRHS[(nz - 1) * st + c] -= (one + eps) * ldn;
        a.gph[x] += (one - eps) * ldn;
        a.gph[(nz - 1) * st + c] -= (one - eps) * ldn;
        LT2[(nz - 1) * st + c] -= dt * g * R(2.0) * d.rdnw[nz - 1] * c2a_at(nz - 1)
                                  * d.alt[(nz - 1) * st + c] * la;
        lmuave += -dt * g * d.c1f[nz] * la;
    }
    f.s[W_L0][c] = LW[c];                        // lower-boundary w''
    // t2 and muave
    real lM0old = half * (one - eps) * lmuave;
    lM1 += half * (one + eps) * lmuave;
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        const real thr = th_ref(d, k, c);
        const real chs = d.c1h[k] * (mut + M1) + d.c2h[k];
        const real t2 = half * ((one + eps) * f.T1[x] + (one - eps) * d.th_pp[x])
                        / (chs * thr);
        GG[x] += half * (one + eps) / (chs * thr) * LT2[x];
        a.gth[x] = half * (one - eps) / (chs * thr) * LT2[x];
        lM1 += -t2 * d.c1h[k] / chs * LT2[x];
    }
    // rhs reverse
    if (s.top_lid) LRHS[nz * st + c] = R(0.0);
    for (int k = 0; k < nz; ++k) LWD[k * st + c] = R(0.0);
    a.grph[c] = R(0.0);
    for (int k = 1; k <= nz; ++k) {
        const int x = k * st + c;
        a.gph[x] += LRHS[x];
        const real lr2 = LRHS[x] / cfm_at(k);
        a.grph[x] = dt * lr2;
        a.gw[x] += dt * half * g * (one - eps) * lr2;
        if (k < nz) {
            LWD[x] += -dt * d.fnm[k] * lr2;
            LWD[(k - 1) * st + c] += -dt * d.fnp[k] * lr2;
        }
    }
    a.gph[c] += gph0;
    for (int k = 0; k <= nz; ++k)
        LWW[k * st + c] = a.lww ? a.lww[k * st + c] : R(0.0);
    for (int k = 0; k < nz; ++k) {
        const real cwd = half * d.rdnw[k] * (phi_ref(d, k + 1, c) - phi_ref(d, k, c));
        LWW[(k + 1) * st + c] += cwd * LWD[k * st + c];
        LWW[k * st + c] += cwd * LWD[k * st + c];
    }
    // advance_mu_th reverse (vertical parts); GG becomes -dt lambda(theta'')
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        const real lT1 = GG[x];
        a.gth[x] += lT1;
        a.grth[x] = dt * lT1;
        GG[x] = -dt * lT1;
    }
    for (int k = 1; k < nz; ++k) {
        // lambda(wdtn[k]) = rdnw[k-1] gg[k-1] - rdnw[k] gg[k]
        const real lwdtn = d.rdnw[k - 1] * GG[(k - 1) * st + c]
                           - d.rdnw[k] * GG[k * st + c];
        LWW[k * st + c] += (d.fnm[k] * th_ref(d, k, c)
                            + d.fnp[k] * th_ref(d, k - 1, c)) * lwdtn;
    }
    real acc = R(0.0), ldmdt = R(0.0), grmu = R(0.0);
    for (int k = nz - 1; k >= 0; --k) {
        // linc[k] = -sum_{kk=k+1}^{nz-1} lww[kk]
        real linc = R(0.0);
        if (k < nz - 1) {
            acc += LWW[(k + 1) * st + c];
            linc = -acc;
        }
        LDVP[k * st + c] = d.dnw[k] * linc;
        ldmdt += d.dnw[k] * d.c1h[k] * linc;
        grmu += d.dnw[k] * d.c1h[k] * linc;
    }
    ldmdt += dt * lM1;
    grmu += dt * lM1;
    a.grmu[c] = grmu;
    a.gmu[c] = lM1 + lM0old;
    for (int k = 0; k < nz; ++k) LDVP[k * st + c] += d.dnw[k] * ldmdt;
    (void)M0;
}

// lambda of u''_new at u face (j, fc) (fc = 0..nx) and of v''_new at v
// face (g, i): the output adjoint plus what the mass cells on both sides
// took from it (divergence, theta flux, lower boundary).
__device__ __forceinline__ real lam_u1(const Dom& d, const Adj& a,
                                       const Fw& f, int k, int j, int fc,
                                       const real* cf, real rdx)
{
    const int ny = d.ny, nx = d.nx, st = ny * nx;
    const real half = R(0.5);
    const real* GG = f.s[W_GG];
    const real* LDVP = f.s[W_LDVP];
    const real* L0 = f.s[W_L0];
    real v = a.lu[iU(k, j, fc, ny, nx)];
    const int ce = fc - 1;                      // cell east of which fc is (fc >= 1)
    const int cw = fc;                          // cell whose west face fc is (fc < nx)
    if (fc >= 1) {
        const int c = j * nx + ce;
        const int cp = j * nx + (fc == nx ? 0 : fc);
        v += half * rdx * (th_ref(d, k, cp) + th_ref(d, k, c)) * GG[k * st + c];
        v += rdx * LDVP[k * st + c];
        if (k < 3) v += cf[k] * half * rdx * (d.ht[cp] - d.ht[c]) * L0[c];
    }
    if (fc < nx) {
        const int c = j * nx + cw;
        const int cm = j * nx + (fc == 0 ? nx - 1 : fc - 1);
        v -= half * rdx * (th_ref(d, k, c) + th_ref(d, k, cm)) * GG[k * st + c];
        v -= rdx * LDVP[k * st + c];
        if (k < 3) v += cf[k] * half * rdx * (d.ht[c] - d.ht[cm]) * L0[c];
    }
    return v;
}

__device__ __forceinline__ real lam_v1(const Dom& d, const Adj& a,
                                       const Fw& f, int k, int g, int i,
                                       const real* cf, real rdy)
{
    const int ny = d.ny, nx = d.nx, st = ny * nx;
    const real half = R(0.5);
    const real* GG = f.s[W_GG];
    const real* LDVP = f.s[W_LDVP];
    const real* L0 = f.s[W_L0];
    real v = a.lv[iV(k, g, i, ny, nx)];
    if (g >= 1) {
        const int c = (g - 1) * nx + i;
        const int cp = (g == ny ? 0 : g) * nx + i;
        v += half * rdy * (th_ref(d, k, cp) + th_ref(d, k, c)) * GG[k * st + c];
        v += rdy * LDVP[k * st + c];
        if (k < 3) v += cf[k] * half * rdy * (d.ht[cp] - d.ht[c]) * L0[c];
    }
    if (g < ny) {
        const int c = g * nx + i;
        const int cm = (g == 0 ? ny - 1 : g - 1) * nx + i;
        v -= half * rdy * (th_ref(d, k, c) + th_ref(d, k, cm)) * GG[k * st + c];
        v -= rdy * LDVP[k * st + c];
        if (k < 3) v += cf[k] * half * rdy * (d.ht[c] - d.ht[cm]) * L0[c];
    }
    return v;
}

// Face pass: u faces (j, fc), fc = 0..nx, then v faces (g, i), g = 0..ny.
// For the non-duplicate faces also the advance_uv transpose pieces: the
// adjoint of the bracket (LB), of the face-averaged damped p'' levels
// folded onto the p'' levels (LS), and of the face mass sum (LM).
__global__ void k_adj_faces(Dom d, Sc s, Adj a, Fw f)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = ny * (nx + 1), nv = (ny + 1) * nx;
    if (t >= nu + nv) return;
    const real half = R(0.5), dt = s.dtau;
    const real cf[3] = {d.cf1, d.cf2, d.cf3};
    const bool isu = t < nu;
    int j, fc, g, i;
    if (isu) { j = t / (nx + 1); fc = t - j * (nx + 1); g = 0; i = 0; }
    else { const int r = t - nu; g = r / nx; i = r - g * nx; j = 0; fc = 0; }
    real* LB = f.s[isu ? W_LBX : W_LBY];
    real* LS = f.s[isu ? W_LSX : W_LSY];
    real* LT = f.s[isu ? W_LPH1 : W_LW];         // dead column scratch reused
    real* LM = f.s[isu ? W_LMX : W_LMY];
    const real rd = isu ? s.rdx : s.rdy;
    // cells A (this face's own index) and B (the one before it)
    int cA, cB;
    bool own;                                    // has its own dpxy
    if (isu) {
        own = fc < nx;
        cA = j * nx + (fc < nx ? fc : 0);
        cB = j * nx + (fc == 0 || fc == nx ? nx - 1 : fc - 1);
    } else {
        own = g < ny;
        cA = (g < ny ? g : 0) * nx + i;
        cB = (g == 0 || g == ny ? ny - 1 : g - 1) * nx + i;
    }
    const int slot = isu ? j * nx + (own ? fc : 0) : (own ? g : 0) * nx + i;
    const real mA = d.mub2d[cA] + d.mup[cA], mB = d.mub2d[cB] + d.mup[cB];
    const real muf = half * (mB + mA);
    real lms = R(0.0);
    for (int k = 0; k < nz; ++k) {
        real l1, ldup = R(0.0);
        if (isu) {
            l1 = lam_u1(d, a, f, k, j, fc, cf, s.rdx);
            a.gu[iU(k, j, fc, ny, nx)] = l1;
            a.gru[iU(k, j, fc, ny, nx)] = dt * l1;
            if (fc == 0) ldup = lam_u1(d, a, f, k, j, nx, cf, s.rdx);
        } else {
            l1 = lam_v1(d, a, f, k, g, i, cf, s.rdy);
            a.gv[iV(k, g, i, ny, nx)] = l1;
            a.grv[iV(k, g, i, ny, nx)] = dt * l1;
            if (g == 0) ldup = lam_v1(d, a, f, k, ny, i, cf, s.rdy);
        }
        if (!own) continue;
        const real ld = -dt * l1 - dt * ldup;
        const real cAk = half * rd * (d.c1h[k] * muf + d.c2h[k]);
        LB[k * st + slot] = cAk * ld;
        real dphp = half * ((d.php[k * st + cA] + d.php[(k + 1) * st + cA])
                            - (d.php[k * st + cB] + d.php[(k + 1) * st + cB]));
        if (d.base3d)
            dphp = dphp + half * ((d.phb[k * st + cA] + d.phb[(k + 1) * st + cA])
                                  - (d.phb[k * st + cB] + d.phb[(k + 1) * st + cB]));
        const real tk = rd * dphp * ld;
        LT[k * st + slot] = tk;
        lms += -half * d.c1h[k] * tk;
    }
    if (!own) return;
    LM[slot] = lms;
    // lambda(dpn) at full levels kf = 0..nz, folded onto the levels k of S
    auto ldpn = [&](int kf) -> real {
        real v = R(0.0);
        if (kf >= 1) v += d.rdnw[kf - 1] * LT[(kf - 1) * st + slot];
        if (kf <= nz - 1) v -= d.rdnw[kf] * LT[kf * st + slot];
        return v;
    };
    const real l0 = ldpn(0), ltop = s.top_lid ? ldpn(nz) : R(0.0);
    for (int k = 0; k < nz; ++k) {
        real v = R(0.0);
        if (k >= 1) v += half * d.fnm[k] * ldpn(k);
        if (k + 1 <= nz - 1) v += half * d.fnp[k + 1] * ldpn(k + 1);
        if (k < 3) v += half * cf[k] * l0;
        if (s.top_lid && k >= nz - 3) v += half * cf[nz - 1 - k] * ltop;
        LS[k * st + slot] = v;
    }
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
    a.gmu[c] += f.s[W_LMX][c] + f.s[W_LMX][ip] + f.s[W_LMY][c] + f.s[W_LMY][jp];
    for (int k = 0; k < nz; ++k) {
        const int x = k * st;
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
    hipLaunchKernelGGL(k_adj_col, dim3((cols + 127) / 128), dim3(128), 0, st,
                       *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_col"))) return rc;
    const int nf = ny * (nx + 1) + (ny + 1) * nx;
    hipLaunchKernelGGL(k_adj_faces, dim3((nf + 127) / 128), dim3(128), 0, st,
                       *d, s, a, fw);
    if ((rc = lib_launched(hipGetLastError(), "k_adj_faces"))) return rc;
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
