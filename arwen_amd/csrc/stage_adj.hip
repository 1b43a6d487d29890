// Hand-written adjoint of the large-step tendency pass of one RK stage
// (dry, periodic, msf = 1, km_opt = 0): the stage-start EOS (reference
// diagnostics.cu:37-89), the coupled mass fluxes and Omega (dycore.py:
// 102-206), the flux-form advection of theta, u, v, w (advection.cu:
// 206-1051), the large-step PGF (dycore.cu:98-211), buoyancy (:231-308)
// and rhs_ph (:541-667).  Given the adjoints of ru_t, rv_t, rw_t, rth_t,
// rph_t it returns those of the stage state u, v, w, thp, php, mu'.
//
// Passes (every kernel gathers, one thread per point it writes):
//   recompute  EOS + Omega (k_omega) and ru, rv (k_faces) into the work
//              area, and theta = thb + thp; the state is not modified
//   advection  the four operator adjoints of advect_adj.hip, adding into
//              the field gradients and the ru / rv / Omega gradients
//   k_sa_pgf   per u / v face and level: the PGF's face-local adjoint
//              pieces (bracket, phi' difference, dpn and mass weights)
//   k_sa_pts   per point: the PGF pieces gathered into p, alpha,
//              alpha_total, php and the mass; buoyancy; rhs_ph (Omega
//              term, g w term, horizontal phi advection)
//   k_sa_omega per column: the transpose of the Omega recurrence
//   k_sa_flux  per face: Omega's divergence and the coupling into u, v
//              (and the mass, parked per face)
//   k_sa_eos   per mass point: the coupling's mass share and the EOS
//   k_sa_php, k_sa_mu  the EOS's phi' share; the column sum into mu'
//
// Also the transposes of the small-step initialisation (through the stage
// EOS) and of the small-step finish, per point with a column pass for the
// column masses.
#include "common.h"
#include "march.h"

namespace ARW_NS {

namespace {

enum {
    S_P, S_AL, S_ALT, S_WW, S_DM, S_RMU, S_RU, S_RV, S_TH,      // recompute
    S_GRU, S_GRV, S_GWW,                                         // flux grads
    S_UBR, S_UDP, S_UHI, S_UMU, S_VBR, S_VDP, S_VHI, S_VMU,      // PGF faces
    S_GPP, S_GAL, S_GALT, S_GMU3, S_GDIV, S_PU, S_PV, S_GDPH,    // gathers
    S_COUNT
};

struct SA {
    Dom d;                   // state, diagnostics pointing into the work area
    const real *lru, *lrv, *lrw, *lrth, *lrph;
    real *gu, *gv, *gw, *gthp, *gphp, *gmup;
    real *gqv, *gqc, *gqr;   // moist: the stage scalars' gradients
    real* w[S_COUNT];
    real rdx, rdy, hrdx, hrdy, qrdx, qrdy;
    int top_lid;
    int acc;                 // the outputs hold values to add to
};

__device__ __forceinline__ int wrapn(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

__device__ __forceinline__ real bv(const real* b, int k, int c, int st,
                                   int b3)
{
    return b3 ? b[k * st + c] : b[k];
}

// p' = p - pb at mass level k, column c
__device__ __forceinline__ real ppv(const SA& a, int k, int c)
{
    const int st = a.d.ny * a.d.nx;
    return a.w[S_P][k * st + c] - bv(a.d.pb, k, c, st, a.d.base3d);
}

// WRF dpn (the perturbation pressure at w level kf) of column c
__device__ real dpn_at(const SA& a, int kf, int c)
{
    const Dom& d = a.d;
    const int nz = d.nz;
    if (kf == 0)
        return (d.cf1 * ppv(a, 0, c) + d.cf2 * ppv(a, 1, c))
               + d.cf3 * ppv(a, 2, c);
    if (kf < nz) return d.fnm[kf] * ppv(a, kf, c) + d.fnp[kf] * ppv(a, kf - 1, c);
    if (a.top_lid) return d.cfn * ppv(a, nz - 1, c) + d.cfn1 * ppv(a, nz - 2, c);
    return R(0.0);
}

__device__ __forceinline__ real phh_at(const SA& a, int k, int c)
{
    const Dom& d = a.d;
    const int st = d.ny * d.nx;
    const real s = d.php[k * st + c] + d.php[(k + 1) * st + c];
    if (d.base3d)
        return R(0.5) * ((d.phb[k * st + c] + d.phb[(k + 1) * st + c]) + s);
    return R(0.5) * s;
}

__device__ __forceinline__ real mut_c(const Dom& d, int c)
{
    return d.mub2d[c] + d.mup[c];
}

// The PGF's face-local adjoint pieces at level k for the face between
// columns A and B (the pressure-gradient term of dycore.cu:98-157 as
// restated in slow_pgf): lt is minus the (folded) adjoint of the face's
// tendency.
__device__ void pgf_pieces(const SA& a, int k, int A, int B, real lt,
                           real rd, real hrd, real& g_br, real& g_dphp,
                           real& g_hi, real& g_mu)
{
    const Dom& d = a.d;
    const int st = d.ny * d.nx;
    const real c1 = d.c1h[k], c2 = d.c2h[k];
    const real muf = R(0.5) * (mut_c(d, A) + mut_c(d, B));
    const real dmu = R(0.5) * (d.mup[A] + d.mup[B]);
    const real layer = c1 * muf + c2;
    const int h = k * st, hp = (k + 1) * st;
    real bracket = (d.php[hp + A] - d.php[hp + B]) + (d.php[h + A] - d.php[h + B]);
    bracket = bracket + (a.w[S_ALT][h + A] + a.w[S_ALT][h + B])
                            * (ppv(a, k, A) - ppv(a, k, B));
    if (d.base3d)
        bracket = bracket + (a.w[S_AL][h + A] + a.w[S_AL][h + B])
                                * (d.pb[h + A] - d.pb[h + B]);
    const real dphp = phh_at(a, k, A) - phh_at(a, k, B);
    const real hi = R(0.5) * (dpn_at(a, k + 1, A) + dpn_at(a, k + 1, B));
    const real lo = R(0.5) * (dpn_at(a, k, A) + dpn_at(a, k, B));
    const real vertical = d.rdnw[k] * (hi - lo) - c1 * dmu;
    g_br = lt * hrd * layer;
    g_dphp = lt * rd * vertical;
    g_hi = lt * rd * dphp * d.rdnw[k];
    g_mu = c1 * lt * (hrd * bracket - rd * dphp);
}

__global__ void k_sa_theta(SA a)
{
    const Dom& d = a.d;
    const int st = d.ny * d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.nz * st) return;
    const int k = t / st, c = t - k * st;
    a.w[S_TH][t] = bv(d.thb, k, c, st, d.base3d) + d.thp[t];
}

// per face (k, j, i), k < nz: u face (j, i) between columns i-1 and i, and
// v face (j, i) between rows j-1 and j
__global__ void k_sa_pgf(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int c = j * nx + i, x = k * ny * nx + c;
    const int cW = j * nx + wrapn(i - 1, nx), cS = wrapn(j - 1, ny) * nx + i;
    real lu = a.lru[iU(k, j, i, ny, nx)];
    if (i == 0) lu += a.lru[iU(k, j, nx, ny, nx)];
    real lv = a.lrv[iV(k, j, i, ny, nx)];
    if (j == 0) lv += a.lrv[iV(k, ny, i, ny, nx)];
    real br, dp, hi, mu;
    pgf_pieces(a, k, c, cW, -lu, a.rdx, a.hrdx, br, dp, hi, mu);
    a.w[S_UBR][x] = br;
    a.w[S_UDP][x] = dp;
    a.w[S_UHI][x] = hi;
    a.w[S_UMU][x] = mu;
    pgf_pieces(a, k, c, cS, -lv, a.rdy, a.hrdy, br, dp, hi, mu);
    a.w[S_VBR][x] = br;
    a.w[S_VDP][x] = dp;
    a.w[S_VHI][x] = hi;
    a.w[S_VMU][x] = mu;
}

// adjoint of dpn at w level kf of column c (E / N: the east / north
// neighbour column, the B side of the faces east and north of c)
__device__ real g_dpn(const SA& a, int kf, int c, int cE, int cN)
{
    const int nz = a.d.nz, st = a.d.ny * a.d.nx;
    real g = R(0.0);
    if (kf >= 1) {
        const int h = (kf - 1) * st;
        g += R(0.5) * (a.w[S_UHI][h + c] + a.w[S_UHI][h + cE])
             + R(0.5) * (a.w[S_VHI][h + c] + a.w[S_VHI][h + cN]);
    }
    if (kf <= nz - 1) {
        const int h = kf * st;
        g -= R(0.5) * (a.w[S_UHI][h + c] + a.w[S_UHI][h + cE])
             + R(0.5) * (a.w[S_VHI][h + c] + a.w[S_VHI][h + cN]);
    }
    return g;
}

// rhs_ph's horizontal term at w level k (1..nz) for the x face i (between
// columns i-1 and i) of row j, or the y face: its adjoint g_f and the
// forward factors m (mass), wv (momentum) and dp (phi difference)
struct GeoFace {
    real gf, m, wv, dp;
};

__device__ GeoFace geo_face(const SA& a, int k, int j, int i, bool xdir)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const bool top = (k == nz);
    const int c = j * nx + i;
    const int cb = xdir ? j * nx + wrapn(i - 1, nx) : wrapn(j - 1, ny) * nx + i;
    const real q = (top ? R(2.0) : R(1.0)) * (xdir ? a.qrdx : a.qrdy);
    GeoFace f;
    const real* lph = a.lrph;
    f.gf = -q * (lph[k * st + c] + lph[k * st + cb]);
    const real avg = R(0.5) * (mut_c(d, c) + mut_c(d, cb));
    f.m = d.c1f[k] * avg + d.c2f[k];
    auto mom = [&](int kk) {
        return xdir ? d.u[iU(kk, j, i, ny, nx)] : d.v[iV(kk, j, i, ny, nx)];
    };
    f.wv = top ? d.cfn * mom(nz - 1) + d.cfn1 * mom(nz - 2)
               : mom(k) + mom(k - 1);
    const real pc = bv(d.phb, k, c, st, d.base3d) + d.php[k * st + c];
    const real pb = bv(d.phb, k, cb, st, d.base3d) + d.php[k * st + cb];
    f.dp = pc - pb;
    return f;
}

// adjoint of rhs_ph's Omega term at mass level m (0..nz-1): the wd[m]
// weight
__device__ __forceinline__ real g_wd(const SA& a, int m, int c)
{
    const Dom& d = a.d;
    const int nz = d.nz, st = d.ny * d.nx;
    real g = R(0.0);
    if (m >= 1 && m <= nz - 1) g -= d.fnm[m] * a.lrph[m * st + c];
    if (m + 1 <= nz - 1) g -= d.fnp[m + 1] * a.lrph[(m + 1) * st + c];
    return g;
}

__device__ __forceinline__ real dph_full(const Dom& d, int m, int c)
{
    const int st = d.ny * d.nx;
    return (bv(d.phb, m + 1, c, st, d.base3d) + d.php[(m + 1) * st + c])
           - (bv(d.phb, m, c, st, d.base3d) + d.php[m * st + c]);
}

// per point of the union grid: mass point (k < nz) gradients of p', alpha,
// alpha_total; w level k gradients of php, w, Omega; u / v points of rhs_ph;
// the mass contributions of all of them into S_GMU3 at (k, c)
__global__ void k_sa_pts(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    const real G = G_F;
    if (i < nx && j < ny) {
        const int c = j * nx + i;
        const int cE = j * nx + wrapn(i + 1, nx), cW = j * nx + wrapn(i - 1, nx);
        const int cN = wrapn(j + 1, ny) * nx + i, cS = wrapn(j - 1, ny) * nx + i;
        real gmu = R(0.0);
        if (k < nz) {
            const int h = k * st;
            const real* alt = a.w[S_ALT];
            const real* UBR = a.w[S_UBR];
            const real* VBR = a.w[S_VBR];
            // PGF bracket: (alt_A + alt_B)(pp_A - pp_B), (al_A + al_B)(pb_A - pb_B)
            real gpp = UBR[h + c] * (alt[h + c] + alt[h + cW])
                       - UBR[h + cE] * (alt[h + cE] + alt[h + c])
                       + VBR[h + c] * (alt[h + c] + alt[h + cS])
                       - VBR[h + cN] * (alt[h + cN] + alt[h + c]);
            const real pc = ppv(a, k, c);
            real galt = UBR[h + c] * (pc - ppv(a, k, cW))
                        + UBR[h + cE] * (ppv(a, k, cE) - pc)
                        + VBR[h + c] * (pc - ppv(a, k, cS))
                        + VBR[h + cN] * (ppv(a, k, cN) - pc);
            real gal = R(0.0);
            if (d.base3d) {
                const real* pb = d.pb;
                gal = UBR[h + c] * (pb[h + c] - pb[h + cW])
                      + UBR[h + cE] * (pb[h + cE] - pb[h + c])
                      + VBR[h + c] * (pb[h + c] - pb[h + cS])
                      + VBR[h + cN] * (pb[h + cN] - pb[h + c]);
            }
            // dpn levels reading p' at level k
            if (k >= 1) gpp += d.fnm[k] * g_dpn(a, k, c, cE, cN);
            if (k + 1 <= nz - 1) gpp += d.fnp[k + 1] * g_dpn(a, k + 1, c, cE, cN);
            if (k <= 2) {
                const real cf = (k == 0) ? d.cf1 : (k == 1 ? d.cf2 : d.cf3);
                gpp += cf * g_dpn(a, 0, c, cE, cN);
            }
            if (a.top_lid) {
                if (k == nz - 1) gpp += d.cfn * g_dpn(a, nz, c, cE, cN);
                if (k == nz - 2) gpp += d.cfn1 * g_dpn(a, nz, c, cE, cN);
            }
            // buoyancy
            if (!d.qv) {
                if (k >= 1) gpp += G * d.rdn[k] * a.lrw[h + c];
                if (k + 1 <= nz - 1) gpp -= G * d.rdn[k + 1] * a.lrw[h + st + c];
                if (k == nz - 1) gpp -= G * R(2.0) * d.rdnw[nz - 1] * a.lrw[nz * st + c];
            } else {
                // with the moist loading (dycore.cu:289-300): the p' term
                // scaled by cq1 = 1 / (1 + cqw), cqw the face mean of
                // qv + qc + qr, and d(term)/d(cqw) = -cq1^2 (rdn dp' + c1f
                // mub + c2f); each face mean takes half of it
                auto qt = [&](int kk) {
                    const int y = kk * st + c;
                    return (d.qv[y] + d.qc[y]) + d.qr[y];
                };
                const real mub = d.mub2d[c];
                real gqt = R(0.0);
                if (k >= 1) {
                    const real q1 = R(1.0) / (R(1.0) + R(0.5) * (qt(k) + qt(k - 1)));
                    const real lw = a.lrw[h + c];
                    gpp += G * q1 * d.rdn[k] * lw;
                    const real dp = ppv(a, k, c) - ppv(a, k - 1, c);
                    gqt += R(0.5) * (G * lw * (-q1 * q1)
                                     * (d.rdn[k] * dp + (d.c1f[k] * mub + d.c2f[k])));
                }
                if (k + 1 <= nz - 1) {
                    const real q1 = R(1.0) / (R(1.0) + R(0.5) * (qt(k + 1) + qt(k)));
                    const real lw = a.lrw[h + st + c];
                    gpp -= G * q1 * d.rdn[k + 1] * lw;
                    const real dp = ppv(a, k + 1, c) - ppv(a, k, c);
                    gqt += R(0.5) * (G * lw * (-q1 * q1)
                                     * (d.rdn[k + 1] * dp + (d.c1f[k + 1] * mub + d.c2f[k + 1])));
                }
                if (k >= nz - 2) {
                    const real q1 = R(1.0) / (R(1.0) + R(0.5) * (qt(nz - 1) + qt(nz - 2)));
                    const real lw = a.lrw[nz * st + c];
                    if (k == nz - 1) gpp -= G * q1 * R(2.0) * d.rdnw[nz - 1] * lw;
                    gqt += R(0.5) * (G * lw * (-q1 * q1)
                                     * (R(2.0) * d.rdnw[nz - 1] * (R(0.0) - ppv(a, nz - 1, c))
                                        + (d.c1f[nz] * mub + d.c2f[nz])));
                }
                a.gqv[h + c] = (a.acc ? a.gqv[h + c] : R(0.0)) + gqt;
                a.gqc[h + c] = (a.acc ? a.gqc[h + c] : R(0.0)) + gqt;
                a.gqr[h + c] = (a.acc ? a.gqr[h + c] : R(0.0)) + gqt;
            }
            a.w[S_GPP][h + c] = gpp;
            a.w[S_GALT][h + c] = galt;
            a.w[S_GAL][h + c] = gal;
            gmu += R(0.5) * (a.w[S_UMU][h + c] + a.w[S_UMU][h + cE])
                   + R(0.5) * (a.w[S_VMU][h + c] + a.w[S_VMU][h + cN]);
        }
        // ---- w level k ----
        const int h = k * st;
        if (k >= 1 && k <= nz - 1) gmu -= G * d.c1f[k] * a.lrw[h + c];
        if (k == nz) gmu -= G * d.c1f[nz] * a.lrw[h + c];
        real gphp = R(0.0), gw = R(0.0), gww = R(0.0);
        const real mt = mut_c(d, c);
        if (k >= 1) {
            gmu += d.c1f[k] * G * d.w[h + c] * a.lrph[h + c];
            gw += (d.c1f[k] * mt + d.c2f[k]) * G * a.lrph[h + c];
        }
        // rhs_ph Omega term
        for (int m = k - 1; m <= k; ++m) {
            if (m < 0 || m > nz - 1) continue;
            const real gd = g_wd(a, m, c);
            const real wavg = R(0.5) * (a.w[S_WW][(m + 1) * st + c]
                                        + a.w[S_WW][m * st + c]);
            gww += R(0.5) * d.rdnw[m] * dph_full(d, m, c) * gd;
            const real t = gd * wavg * d.rdnw[m];
            gphp += (m == k - 1) ? t : -t;
        }
        // rhs_ph horizontal term: faces i, i+1 (x) and j, j+1 (y)
        if (k >= 1) {
            const GeoFace fx0 = geo_face(a, k, j, i, true);
            const GeoFace fx1 = geo_face(a, k, j, wrapn(i + 1, nx), true);
            const GeoFace fy0 = geo_face(a, k, j, i, false);
            const GeoFace fy1 = geo_face(a, k, wrapn(j + 1, ny), i, false);
            gphp += (fx0.gf * fx0.m * fx0.wv - fx1.gf * fx1.m * fx1.wv)
                    + (fy0.gf * fy0.m * fy0.wv - fy1.gf * fy1.m * fy1.wv);
            const real c1 = d.c1f[k];
            gmu += R(0.5) * c1 * ((fx0.gf * fx0.dp * fx0.wv + fx1.gf * fx1.dp * fx1.wv)
                                  + (fy0.gf * fy0.dp * fy0.wv + fy1.gf * fy1.dp * fy1.wv));
        }
        // PGF php terms: bracket at levels k-1, k; phh at levels k-1, k
        for (int kl = k - 1; kl <= k; ++kl) {
            if (kl < 0 || kl > nz - 1) continue;
            const int hl = kl * st;
            gphp += (a.w[S_UBR][hl + c] - a.w[S_UBR][hl + cE])
                    + (a.w[S_VBR][hl + c] - a.w[S_VBR][hl + cN]);
            gphp += R(0.5) * ((a.w[S_UDP][hl + c] - a.w[S_UDP][hl + cE])
                              + (a.w[S_VDP][hl + c] - a.w[S_VDP][hl + cN]));
        }
        a.gphp[h + c] += gphp;
        a.gw[h + c] += gw;
        a.w[S_GWW][h + c] += gww;
        a.w[S_GMU3][h + c] = gmu;
    }
    // u point (k, j, i), i < nx: rhs_ph's horizontal term reads u at w
    // levels k (as u[kk]) and k+1 (as u[kk-1]), and the lid extrapolation
    if (k < nz && j < ny && i < nx) {
        real g = R(0.0);
        auto gwv = [&](int kk) {
            const GeoFace f = geo_face(a, kk, j, i, true);
            return f.gf * f.m * f.dp;
        };
        if (k >= 1) g += gwv(k);
        if (k + 1 <= nz - 1) g += gwv(k + 1);
        if (k == nz - 1) g += d.cfn * gwv(nz);
        if (k == nz - 2) g += d.cfn1 * gwv(nz);
        a.gu[iU(k, j, i, ny, nx)] += g;
    }
    if (k < nz && i < nx && j < ny) {
        real g = R(0.0);
        auto gwv = [&](int kk) {
            const GeoFace f = geo_face(a, kk, j, i, false);
            return f.gf * f.m * f.dp;
        };
        if (k >= 1) g += gwv(k);
        if (k + 1 <= nz - 1) g += gwv(k + 1);
        if (k == nz - 1) g += d.cfn * gwv(nz);
        if (k == nz - 2) g += d.cfn1 * gwv(nz);
        a.gv[iV(k, j, i, ny, nx)] += g;
    }
}

// Omega: ww[k+1] = -sum_{m<=k} ((c1h dnw)[m] dmdt + divv[m]), dmdt =
// sum_m divv[m]; transpose per column
__global__ void k_sa_omega(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, st = d.ny * d.nx;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const real* gww = a.w[S_GWW];
    real* gdiv = a.w[S_GDIV];
    real run = R(0.0), gdm = R(0.0);
    for (int k = nz - 2; k >= 0; --k) {
        run = run - gww[(k + 1) * st + c];
        gdiv[k * st + c] = run;                 // g_inc, for now
        gdm += (d.c1h[k] * d.dnw[k]) * run;
    }
    for (int k = 0; k < nz; ++k)
        gdiv[k * st + c] = (k <= nz - 2 ? gdiv[k * st + c] : R(0.0)) + gdm;
}

// per face: Omega's divergence adjoint into ru / rv, then the coupling
// ru = (c1h mu_x + c2h) u into u (and, parked per face, into the mass)
__global__ void k_sa_flux(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k >= nz) return;
    const real* gdiv = a.w[S_GDIV];
    const int h = k * st;
    if (j < ny) {                               // u face (k, j, i)
        const int x = iU(k, j, i, ny, nx);
        real g = a.w[S_GRU][x];
        real dv = R(0.0);
        if (i >= 1) dv += gdiv[h + j * nx + i - 1];
        if (i <= nx - 1) dv -= gdiv[h + j * nx + i];
        g += d.dnw[k] * a.rdx * dv;
        const int ca = j * nx + (i >= nx ? i - nx : i);
        const int cb = j * nx + (i == 0 ? nx - 1 : i - 1);
        const real mux = R(0.5) * (mut_c(d, ca) + mut_c(d, cb));
        a.gu[x] += (d.c1h[k] * mux + d.c2h[k]) * g;
        a.w[S_PU][x] = d.c1h[k] * d.u[x] * g;
    }
    if (i < nx) {                               // v face (k, j, i)
        const int x = iV(k, j, i, ny, nx);
        real g = a.w[S_GRV][x];
        real dv = R(0.0);
        if (j >= 1) dv += gdiv[h + (j - 1) * nx + i];
        if (j <= ny - 1) dv -= gdiv[h + j * nx + i];
        g += d.dnw[k] * a.rdy * dv;
        const int ca = (j >= ny ? j - ny : j) * nx + i;
        const int cb = (j == 0 ? ny - 1 : j - 1) * nx + i;
        const real muy = R(0.5) * (mut_c(d, ca) + mut_c(d, cb));
        a.gv[x] += (d.c1h[k] * muy + d.c2h[k]) * g;
        a.w[S_PV][x] = d.c1h[k] * d.v[x] * g;
    }
}

// per mass point: the coupling's mass share (faces folded onto the
// periodic ones) and the EOS transpose (theta', the phi' layer, mu')
__global__ void k_sa_eos(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nz * st) return;
    const int k = t / st, c = t - k * st;
    const int j = c / nx, i = c - j * nx;
    const real* PU = a.w[S_PU];
    const real* PV = a.w[S_PV];
    auto pu = [&](int f) {        // folded: the duplicate face nx adds to 0
        real v = PU[iU(k, j, f, ny, nx)];
        if (f == 0) v += PU[iU(k, j, nx, ny, nx)];
        return v;
    };
    auto pv = [&](int g) {
        real v = PV[iV(k, g, i, ny, nx)];
        if (g == 0) v += PV[iV(k, ny, i, ny, nx)];
        return v;
    };
    real gmu = R(0.5) * (pu(i) + pu(wrapn(i + 1, nx)))
               + R(0.5) * (pv(j) + pv(wrapn(j + 1, ny)));
    const real p = a.w[S_P][t], alt = a.w[S_ALT][t];
    const real th = bv(d.thb, k, c, st, d.base3d) + d.thp[t];
    const real mh = d.c1h[k] * mut_c(d, c) + d.c2h[k];
    const real gp = a.w[S_GPP][t];
    const real ga = a.w[S_GALT][t] + a.w[S_GAL][t] - gp * (GAMMA_F * p / alt);
    a.gthp[t] += gp * (GAMMA_F * p / th);
    if (d.qv) {            // theta_m = theta (1 + Rv/Rd qv) of the moist EOS
        const real thm = th * (R(1.0) + RVOVRD_F * d.qv[t]);
        a.gqv[t] += gp * (GAMMA_F * p / thm) * th * RVOVRD_F;
    }
    a.w[S_GDPH][t] = ga * (-d.rdnw[k] / mh);
    gmu += ga * (-alt * d.c1h[k] / mh);
    a.w[S_GMU3][t] += gmu;
}

// Constant-K diffusion (reference diffusion.cu:22-51, coupled in
// diffusion.py:100-130; flat base): tend += C(mu_face) diff2(field), with
// diff2 the Laplacian kh (fxx + fyy) + kv (fz)z on each field's periodic
// core (the duplicate face copies the first; w's end levels are zero).
// One thread per cell and w level gathers each field's transpose at its
// core point -- the horizontal 3-point operator is symmetric, the vertical
// one maps mu_k = kv L_k rdzc_k to rdzf_(k-1) (mu_(k-1) - mu_k) + rdzf_k
// (mu_(k+1) - mu_k) -- and adds the column-mass share C'(mu) diff2 lam
// (the u / v faces' halved onto their two cells) to the level's mass piece.
struct DiffField {
    const real* f;       // the field
    const real* lam;     // adjoint of its tendency
    const real* rdzf;    // inverse spacings of its levels
    const real* rdzc;
    int nlev, zero_ends;
};

__device__ __forceinline__ real mu_at(const Dom& d, int jj, int ii)
{
    const int ny = d.ny, nx = d.nx;
    jj = jj < 0 ? jj + ny : (jj >= ny ? jj - ny : jj);
    ii = ii < 0 ? ii + nx : (ii >= nx ? ii - nx : ii);
    return mut_c(d, jj * nx + ii);
}

__global__ void k_sa_diff(SA a, Sc s)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k > nz) return;
    const real half = R(0.5), two = R(2.0);
    auto wi = [&](int q) { return q < 0 ? q + nx : (q >= nx ? q - nx : q); };
    auto wj = [&](int q) { return q < 0 ? q + ny : (q >= ny ? q - ny : q); };
    // face masses: u face f between cells f-1 and f of row jj, v face g
    // between rows g-1 and g of column ii (periodic; face nx / ny = face 0)
    auto mux = [&](int jj, int f) {
        return half * (mu_at(d, jj, f) + mu_at(d, jj, f - 1));
    };
    auto muy = [&](int g, int ii) {
        return half * (mu_at(d, g, ii) + mu_at(d, g - 1, ii));
    };
    // field values on the periodic core
    auto U = [&](int kk, int jj, int ii) { return d.u[iU(kk, wj(jj), wi(ii), ny, nx)]; };
    auto V = [&](int kk, int jj, int ii) { return d.v[iV(kk, wj(jj), wi(ii), ny, nx)]; };
    auto M = [&](const real* f, int kk, int jj, int ii) {
        return f[(kk * ny + wj(jj)) * nx + wi(ii)];
    };
    // the forward Laplacian of a field at a core point
    auto lap = [&](auto F, int kk, int jj, int ii, const real* rdzf,
                   const real* rdzc, int nlev, int zero_ends) -> real {
        if (zero_ends && (kk == 0 || kk == nlev - 1)) return R(0.0);
        const real fc = F(kk, jj, ii);
        real r = s.kh * (s.dxi2 * (F(kk, jj, ii - 1) - two * fc + F(kk, jj, ii + 1))
                         + s.dyi2 * (F(kk, jj - 1, ii) - two * fc + F(kk, jj + 1, ii)));
        const real fup = kk <= nlev - 2 ? (F(kk + 1, jj, ii) - fc) * rdzf[kk] : R(0.0);
        const real fdn = kk >= 1 ? (fc - F(kk - 1, jj, ii)) * rdzf[kk - 1] : R(0.0);
        return r + s.kv * (fup - fdn) * rdzc[kk];
    };
    // the transpose at a core point for L(kk, jj, ii) = lambda of the
    // Laplacian (coupled, folded, ends zeroed)
    auto lapT = [&](auto L, int kk, int jj, int ii, const real* rdzf,
                    const real* rdzc, int nlev) -> real {
        const real lc = L(kk, jj, ii);
        real g = s.kh * (s.dxi2 * (L(kk, jj, ii - 1) - two * lc + L(kk, jj, ii + 1))
                         + s.dyi2 * (L(kk, jj - 1, ii) - two * lc + L(kk, jj + 1, ii)));
        const real m0 = s.kv * lc * rdzc[kk];
        if (kk >= 1) g += rdzf[kk - 1] * (s.kv * L(kk - 1, jj, ii) * rdzc[kk - 1] - m0);
        if (kk <= nlev - 2) g += rdzf[kk] * (s.kv * L(kk + 1, jj, ii) * rdzc[kk + 1] - m0);
        return g;
    };
    const int c = j * nx + i;
    real piece = R(0.0);
    // w (levels 0..nz)
    {
        auto LW = [&](int kk, int jj, int ii) -> real {
            if (kk == 0 || kk == nz) return R(0.0);
            const int y = (kk * ny + wj(jj)) * nx + wi(ii);
            return (d.c1f[kk] * mu_at(d, jj, ii) + d.c2f[kk]) * a.lrw[y];
        };
        auto FW = [&](int kk, int jj, int ii) { return M(d.w, kk, jj, ii); };
        a.gw[k * st + c] += lapT(LW, k, j, i, d.rdzf_w, d.rdzc_w, nz + 1);
        piece += d.c1f[k] * lap(FW, k, j, i, d.rdzf_w, d.rdzc_w, nz + 1, 1)
                 * a.lrw[k * st + c];
    }
    if (k < nz) {
        // theta'
        auto LT = [&](int kk, int jj, int ii) -> real {
            const int y = (kk * ny + wj(jj)) * nx + wi(ii);
            return (d.c1h[kk] * mu_at(d, jj, ii) + d.c2h[kk]) * a.lrth[y];
        };
        auto FT = [&](int kk, int jj, int ii) { return M(d.thp, kk, jj, ii); };
        a.gthp[k * st + c] += lapT(LT, k, j, i, d.rdzf_h, d.rdzc_h, nz);
        piece += d.c1h[k] * lap(FT, k, j, i, d.rdzf_h, d.rdzc_h, nz, 0)
                 * a.lrth[k * st + c];
        // u (core faces 0..nx-1; face nx folds into face 0)
        auto LU = [&](int kk, int jj, int ii) -> real {
            const int jr = wj(jj), f = wi(ii);
            real v = (d.c1h[kk] * mux(jr, f) + d.c2h[kk]) * a.lru[iU(kk, jr, f, ny, nx)];
            if (f == 0)
                v += (d.c1h[kk] * mux(jr, nx) + d.c2h[kk]) * a.lru[iU(kk, jr, nx, ny, nx)];
            return v;
        };
        a.gu[iU(k, j, i, ny, nx)] += lapT(LU, k, j, i, d.rdzf_h, d.rdzc_h, nz);
        auto gfu = [&](int f) {        // folded mass share of u face f
            const real t = lap(U, k, j, f, d.rdzf_h, d.rdzc_h, nz, 0);
            real v = d.c1h[k] * t * a.lru[iU(k, j, f, ny, nx)];
            if (f == 0) v += d.c1h[k] * t * a.lru[iU(k, j, nx, ny, nx)];
            return v;
        };
        piece += half * (gfu(i) + gfu(wi(i + 1)));
        // v (core rows 0..ny-1; row ny folds into row 0)
        auto LV = [&](int kk, int jj, int ii) -> real {
            const int g = wj(jj), ic = wi(ii);
            real v = (d.c1h[kk] * muy(g, ic) + d.c2h[kk]) * a.lrv[iV(kk, g, ic, ny, nx)];
            if (g == 0)
                v += (d.c1h[kk] * muy(ny, ic) + d.c2h[kk]) * a.lrv[iV(kk, ny, ic, ny, nx)];
            return v;
        };
        a.gv[iV(k, j, i, ny, nx)] += lapT(LV, k, j, i, d.rdzf_h, d.rdzc_h, nz);
        auto gfv = [&](int g) {
            const real t = lap(V, k, g, i, d.rdzf_h, d.rdzc_h, nz, 0);
            real v = d.c1h[k] * t * a.lrv[iV(k, g, i, ny, nx)];
            if (g == 0) v += d.c1h[k] * t * a.lrv[iV(k, ny, i, ny, nx)];
            return v;
        };
        piece += half * (gfv(j) + gfv(wj(j + 1)));
    }
    a.w[S_GMU3][k * st + c] += piece;
}

__global__ void k_sa_php(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, st = d.ny * d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (nz + 1) * st) return;
    const int k = t / st;
    real g = R(0.0);
    if (k >= 1) g += a.w[S_GDPH][t - st];
    if (k <= nz - 1) g -= a.w[S_GDPH][t];
    a.gphp[t] += g;
}

__global__ void k_sa_mu(SA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, st = d.ny * d.nx;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    real s = R(0.0);
    for (int k = 0; k <= nz; ++k) s += a.w[S_GMU3][k * st + c];
    a.gmup[c] = (a.acc ? a.gmup[c] : R(0.0)) + s;
}

// ---------------------------------------------------------------------------
// Small-step initialisation (dycore.cu:947-1085) and finish (:1087-1198)
// ---------------------------------------------------------------------------

enum {
    I_P, I_AL, I_ALT, I_GP, I_GALT, I_GDPH, I_GDPH2, I_M3, I_S3, I_U3,
    I_U3B, I_MTW, I_MSW, I_PU0, I_PU, I_PV0, I_PV, I_COUNT
};
enum { F_S3, F_N3, F_SW, F_NW, F_PSU, F_PNU, F_PSV, F_PNV, F_COUNT };

struct IA {
    Dom d;                   // p / al / alt point at the recomputed EOS
    const real *lu, *lv, *lw, *lth, *lph, *lmu, *lal, *lp, *lpo;
    real *gu0, *gv0, *gw0, *gthp0, *gphp0, *gmup0;
    real *gu, *gv, *gw, *gthp, *gphp, *gmup;
    real *gqv;               // moist: the stage qv's gradient
    real* w[I_COUNT];
    int acc;                 // the outputs hold values to add to
};

// an init-adjoint output: written, or added to what it holds
__device__ __forceinline__ void put_i(const IA& a, real* g, int x, real v)
{
    g[x] = (a.acc ? g[x] : R(0.0)) + v;
}

__device__ __forceinline__ real mu0_c(const Dom& d, int c)
{
    return d.mub2d[c] + d.mup0[c];
}

// per point of the union grid: the mass, w-level and face parts of the
// transposed initialisation
__global__ void k_ia_pts(IA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    if (i < nx && j < ny) {
        const int c = j * nx + i;
        const real mt = mu0_c(d, c), ms = mut_c(d, c);
        if (k < nz) {
            const int t = k * st + c;
            const real c1 = d.c1h[k], c2 = d.c2h[k];
            const real thb = bv(d.thb, k, c, st, d.base3d);
            const real tht = thb + d.thp0[t], ths = thb + d.thp[t];
            const real cht = c1 * mt + c2, chs = c1 * ms + c2;
            const real mupp = d.mup0[c] - d.mup[c];
            const real dpp = (d.php0[t + st] - d.php[t + st])
                             - (d.php0[t] - d.php[t]);
            const real thpp = cht * tht - chs * ths;
            const real p = a.w[I_P][t], alt = a.w[I_ALT][t];
            const real c2a = GAMMA_F * p / alt;
            const real al = -((alt * (c1 * mupp)) + d.rdnw[k] * dpp) / cht;
            const real N = thpp - (c1 * mupp) * ths;
            const real p5 = (alt * N) / (cht * ths);
            const real Lp = a.lp[t] + a.lpo[t];
            const real g_c2a = Lp * (p5 - al), g_p5 = Lp * c2a;
            const real g_al = a.lal[t] - Lp * c2a;
            real g_alt = g_p5 * N / (cht * ths);
            const real g_N = g_p5 * alt / (cht * ths);
            real g_cht = -g_p5 * p5 / cht;
            real g_ths = -g_p5 * p5 / ths - g_N * (c1 * mupp);
            const real g_thpp = g_N + a.lth[t];
            real g_mupp = -g_N * c1 * ths;
            g_alt = g_alt - g_al * c1 * mupp / cht;
            g_mupp = g_mupp - g_al * alt * c1 / cht;
            const real g_dph = -g_al * d.rdnw[k] / cht;
            g_cht = g_cht - g_al * al / cht;
            const real g_p = g_c2a * GAMMA_F / alt;
            g_alt = g_alt - g_c2a * c2a / alt;
            g_cht = g_cht + g_thpp * tht;
            const real g_chs = -g_thpp * ths;
            g_ths = g_ths - g_thpp * chs;
            put_i(a, a.gthp0, t, g_thpp * cht);
            put_i(a, a.gthp, t, g_ths);
            a.w[I_GP][t] = g_p;
            a.w[I_GALT][t] = g_alt;
            a.w[I_GDPH][t] = g_dph;
            a.w[I_M3][t] = c1 * g_cht;
            a.w[I_S3][t] = c1 * g_chs;
            a.w[I_U3][t] = g_mupp;
        }
        const int t = k * st + c;              // w level k
        const real lw = a.lw[t];
        put_i(a, a.gw0, t, (d.c1f[k] * mt + d.c2f[k]) * lw);
        put_i(a, a.gw, t, -((d.c1f[k] * ms + d.c2f[k]) * lw));
        a.w[I_MTW][t] = d.c1f[k] * d.w0[t] * lw;
        a.w[I_MSW][t] = -(d.c1f[k] * d.w[t] * lw);
    }
    if (k < nz && j < ny) {                    // u face (k, j, i)
        const int x = iU(k, j, i, ny, nx);
        const int ca = j * nx + (i >= nx ? i - nx : i);
        const int cb = j * nx + (i == 0 ? nx - 1 : i - 1);
        const real mtf = R(0.5) * (mu0_c(d, ca) + mu0_c(d, cb));
        const real msf = R(0.5) * (mut_c(d, ca) + mut_c(d, cb));
        const real c1 = d.c1h[k], c2 = d.c2h[k], L = a.lu[x];
        put_i(a, a.gu0, x, (c1 * mtf + c2) * L);
        put_i(a, a.gu, x, -((c1 * msf + c2) * L));
        a.w[I_PU0][x] = c1 * d.u0[x] * L;
        a.w[I_PU][x] = -(c1 * d.u[x] * L);
    }
    if (k < nz && i < nx) {                    // v face (k, j, i)
        const int x = iV(k, j, i, ny, nx);
        const int ca = (j >= ny ? j - ny : j) * nx + i;
        const int cb = (j == 0 ? ny - 1 : j - 1) * nx + i;
        const real mtf = R(0.5) * (mu0_c(d, ca) + mu0_c(d, cb));
        const real msf = R(0.5) * (mut_c(d, ca) + mut_c(d, cb));
        const real c1 = d.c1h[k], c2 = d.c2h[k], L = a.lv[x];
        put_i(a, a.gv0, x, (c1 * mtf + c2) * L);
        put_i(a, a.gv, x, -((c1 * msf + c2) * L));
        a.w[I_PV0][x] = c1 * d.v0[x] * L;
        a.w[I_PV][x] = -(c1 * d.v[x] * L);
    }
}

// the stage EOS's transpose (p and alt gradients of the initialisation)
__global__ void k_ia_eos(IA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, st = d.ny * d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nz * st) return;
    const int k = t / st, c = t - k * st;
    const real p = a.w[I_P][t], alt = a.w[I_ALT][t];
    const real th = bv(d.thb, k, c, st, d.base3d) + d.thp[t];
    const real mh = d.c1h[k] * mut_c(d, c) + d.c2h[k];
    const real gp = a.w[I_GP][t];
    const real ga = a.w[I_GALT][t] - gp * (GAMMA_F * p / alt);
    a.gthp[t] += gp * (GAMMA_F * p / th);
    if (d.qv) {
        const real thm = th * (R(1.0) + RVOVRD_F * d.qv[t]);
        put_i(a, a.gqv, t, gp * (GAMMA_F * p / thm) * th * RVOVRD_F);
    }
    a.w[I_GDPH2][t] = ga * (-d.rdnw[k] / mh);
    a.w[I_U3B][t] = ga * (-alt * d.c1h[k] / mh);
}

// w levels: ph'' = php0 - php through al'', and the EOS's phi' layer
__global__ void k_ia_php(IA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, st = d.ny * d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (nz + 1) * st) return;
    const int k = t / st;
    real g = a.lph[t], e = R(0.0);
    if (k >= 1) {
        g += a.w[I_GDPH][t - st];
        e += a.w[I_GDPH2][t - st];
    }
    if (k <= nz - 1) {
        g -= a.w[I_GDPH][t];
        e -= a.w[I_GDPH2][t];
    }
    put_i(a, a.gphp0, t, g);
    put_i(a, a.gphp, t, -g + e);
}

// the face products of a column folded onto the periodic faces c, c+1
__device__ __forceinline__ real face_pair_u(const real* P, int k, int j,
                                            int i, int ny, int nx)
{
    auto f = [&](int ff) {
        real v = P[iU(k, j, ff, ny, nx)];
        if (ff == 0) v += P[iU(k, j, nx, ny, nx)];
        return v;
    };
    return R(0.5) * (f(i) + f(i + 1 == nx ? 0 : i + 1));
}

__device__ __forceinline__ real face_pair_v(const real* P, int k, int j,
                                            int i, int ny, int nx)
{
    auto f = [&](int gg) {
        real v = P[iV(k, gg, i, ny, nx)];
        if (gg == 0) v += P[iV(k, ny, i, ny, nx)];
        return v;
    };
    return R(0.5) * (f(j) + f(j + 1 == ny ? 0 : j + 1));
}

// per column: the time-t and stage mass gradients
__global__ __launch_bounds__(64 * KG) void k_ia_cols(IA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const bool on = c < st;
    const int cc = on ? c : 0;
    const int j = cc / nx, i = cc - j * nx;
    real v[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};   // gmt, gms, gmupp, geos
    if (on) {
        for (int k = ty; k < nz; k += KG) {
            const int t = k * st + c;
            v[0] += a.w[I_M3][t] + face_pair_u(a.w[I_PU0], k, j, i, ny, nx)
                    + face_pair_v(a.w[I_PV0], k, j, i, ny, nx);
            v[1] += a.w[I_S3][t] + face_pair_u(a.w[I_PU], k, j, i, ny, nx)
                    + face_pair_v(a.w[I_PV], k, j, i, ny, nx);
            v[2] += a.w[I_U3][t];
            v[3] += a.w[I_U3B][t];
        }
        for (int k = ty; k <= nz; k += KG) {
            v[0] += a.w[I_MTW][k * st + c];
            v[1] += a.w[I_MSW][k * st + c];
        }
    }
    group_sums(v);
    if (!on || ty != 0) return;
    put_i(a, a.gmup0, c, (a.lmu[c] + v[2]) + v[0]);
    put_i(a, a.gmup, c, (v[1] - (a.lmu[c] + v[2])) + v[3]);
}

struct FA {
    Dom d;
    const real *lu, *lv, *lw, *lthp, *lphp, *lmup;
    real *gu, *gv, *gw, *gthp, *gphp, *gmup;
    real *gupp, *gvpp, *gwpp, *gthpp, *gphpp, *gmupp;
    real* w[F_COUNT];
};

__global__ void k_fa_pts(FA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i > nx || j > ny || k > nz) return;
    auto mn_c = [&](int c) { return mut_c(d, c) + d.mu_pp[c]; };
    if (k < nz && j < ny) {                    // u face (k, j, i)
        const int x = iU(k, j, i, ny, nx);
        const int ca = j * nx + (i >= nx ? i - nx : i);
        const int cb = j * nx + (i == 0 ? nx - 1 : i - 1);
        const real msf = R(0.5) * (mut_c(d, ca) + mut_c(d, cb));
        const real mnf = R(0.5) * (mn_c(ca) + mn_c(cb));
        const real c1 = d.c1h[k], c2 = d.c2h[k];
        const real den = c1 * mnf + c2;
        const real nw = ((c1 * msf + c2) * d.u[x] + d.u_pp[x]) / den;
        const real gn = a.lu[x] / den;
        a.gu[x] = (c1 * msf + c2) * gn;
        a.gupp[x] = gn;
        a.w[F_PSU][x] = c1 * d.u[x] * gn;
        a.w[F_PNU][x] = -(c1 * gn * nw);
    }
    if (k < nz && i < nx) {                    // v face (k, j, i)
        const int x = iV(k, j, i, ny, nx);
        const int ca = (j >= ny ? j - ny : j) * nx + i;
        const int cb = (j == 0 ? ny - 1 : j - 1) * nx + i;
        const real msf = R(0.5) * (mut_c(d, ca) + mut_c(d, cb));
        const real mnf = R(0.5) * (mn_c(ca) + mn_c(cb));
        const real c1 = d.c1h[k], c2 = d.c2h[k];
        const real den = c1 * mnf + c2;
        const real nw = ((c1 * msf + c2) * d.v[x] + d.v_pp[x]) / den;
        const real gn = a.lv[x] / den;
        a.gv[x] = (c1 * msf + c2) * gn;
        a.gvpp[x] = gn;
        a.w[F_PSV][x] = c1 * d.v[x] * gn;
        a.w[F_PNV][x] = -(c1 * gn * nw);
    }
    if (i < nx && j < ny) {
        const int c = j * nx + i;
        const real ms = mut_c(d, c), mn = ms + d.mu_pp[c];
        const int t = k * st + c;               // w level k
        {
            const real c1 = d.c1f[k], c2 = d.c2f[k];
            const real den = c1 * mn + c2;
            const real nw = ((c1 * ms + c2) * d.w[t] + d.w_pp[t]) / den;
            const real gn = a.lw[t] / den;
            a.gw[t] = (c1 * ms + c2) * gn;
            a.gwpp[t] = gn;
            a.w[F_SW][t] = c1 * d.w[t] * gn;
            a.w[F_NW][t] = -(c1 * gn * nw);
            a.gphp[t] = a.lphp[t];
            a.gphpp[t] = a.lphp[t];
        }
        if (k < nz) {                           // mass level k
            const real c1 = d.c1h[k], c2 = d.c2h[k];
            const real thb = bv(d.thb, k, c, st, d.base3d);
            const real den = c1 * mn + c2;
            const real th = ((c1 * ms + c2) * (thb + d.thp[t]) + d.th_pp[t]) / den;
            const real gn = a.lthp[t] / den;
            a.gthp[t] = (c1 * ms + c2) * gn;
            a.gthpp[t] = gn;
            a.w[F_S3][t] = c1 * (thb + d.thp[t]) * gn;
            a.w[F_N3][t] = -(c1 * gn * th);
        }
    }
}

__global__ __launch_bounds__(64 * KG) void k_fa_cols(FA a)
{
    const Dom& d = a.d;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * 64 + threadIdx.x, ty = threadIdx.y;
    const bool on = c < st;
    const int cc = on ? c : 0;
    const int j = cc / nx, i = cc - j * nx;
    real v[2] = {R(0.0), R(0.0)};                   // gms, gmn
    if (on) {
        for (int k = ty; k < nz; k += KG) {
            v[0] += a.w[F_S3][k * st + c] + face_pair_u(a.w[F_PSU], k, j, i, ny, nx)
                    + face_pair_v(a.w[F_PSV], k, j, i, ny, nx);
            v[1] += a.w[F_N3][k * st + c] + face_pair_u(a.w[F_PNU], k, j, i, ny, nx)
                    + face_pair_v(a.w[F_PNV], k, j, i, ny, nx);
        }
        for (int k = ty; k <= nz; k += KG) {
            v[0] += a.w[F_SW][k * st + c];
            v[1] += a.w[F_NW][k * st + c];
        }
    }
    group_sums(v);
    if (!on || ty != 0) return;
    const real L = a.lmup[c];
    a.gmup[c] = (L + v[0]) + v[1];
    a.gmupp[c] = L + v[1];
}

}  // namespace

}  // namespace ARW_NS

using namespace ARW_NS;

extern "C" int32_t ARW_API(small_step_adjoint_work)(int32_t nz, int32_t ny,
                                                    int32_t nx,
                                                    int64_t* elements)
{
    if (!elements) return lib_fail(ARW_E_ARG, "null argument");
    if (nz < 3 || ny < 1 || nx < 1) return lib_fail(ARW_E_ARG, "bad extents");
    *elements = (int64_t)I_COUNT * (nz + 1) * (ny + 1) * (nx + 1);
    return 0;
}

extern "C" int32_t ARW_API(small_step_init_adjoint)(
    const Dom* d, const ARW_API(init_adj)* adj, real* work, void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
    if (rc) return rc;
    if (!adj || !work) return lib_fail(ARW_E_ARG, "null argument");
    const real* lin[] = {adj->l_u_pp, adj->l_v_pp, adj->l_w_pp, adj->l_th_pp,
                         adj->l_ph_pp, adj->l_mu_pp, adj->l_al_pp,
                         adj->l_p_pp, adj->l_p_pp_old};
    real* gout[] = {adj->g_u0, adj->g_v0, adj->g_w0, adj->g_thp0,
                    adj->g_php0, adj->g_mup0, adj->g_u, adj->g_v, adj->g_w,
                    adj->g_thp, adj->g_php, adj->g_mup};
    for (const real* p : lin)
        if (!p) return lib_fail(ARW_E_ARG, "null output adjoint");
    for (real* p : gout)
        if (!p) return lib_fail(ARW_E_ARG, "null input adjoint");
    if (d->qv && !adj->g_qv)
        return lib_fail(ARW_E_ARG, "init adjoint: a moist state needs g_qv");
    const int nz = d->nz, ny = d->ny, nx = d->nx, st = ny * nx;
    if (nz < 3) return lib_fail(ARW_E_ARG, "init adjoint needs nz >= 3");
    hipStream_t sm = (hipStream_t)stream;
    const size_t F = (size_t)(nz + 1) * (ny + 1) * (nx + 1);
    IA a;
    a.d = *d;
    for (int w = 0; w < I_COUNT; ++w) a.w[w] = work + (size_t)w * F;
    a.lu = adj->l_u_pp; a.lv = adj->l_v_pp; a.lw = adj->l_w_pp;
    a.lth = adj->l_th_pp; a.lph = adj->l_ph_pp; a.lmu = adj->l_mu_pp;
    a.lal = adj->l_al_pp; a.lp = adj->l_p_pp; a.lpo = adj->l_p_pp_old;
    a.gu0 = adj->g_u0; a.gv0 = adj->g_v0; a.gw0 = adj->g_w0;
    a.gthp0 = adj->g_thp0; a.gphp0 = adj->g_php0; a.gmup0 = adj->g_mup0;
    a.gu = adj->g_u; a.gv = adj->g_v; a.gw = adj->g_w; a.gthp = adj->g_thp;
    a.gphp = adj->g_php; a.gmup = adj->g_mup; a.gqv = adj->g_qv;
    if (adj->flags & ~ARW_ADJ_ACCUMULATE)
        return lib_fail(ARW_E_ARG, "unknown adjoint flags");
    a.acc = (adj->flags & ARW_ADJ_ACCUMULATE) != 0;
    a.d.p = a.w[I_P]; a.d.al = a.w[I_AL]; a.d.alt = a.w[I_ALT];
    if ((rc = lib_launched(run_eos(a.d, sm), "k_eos (init adjoint)"))) return rc;
    const dim3 blk(64, 4);
    const dim3 gu((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz + 1);
    hipLaunchKernelGGL(k_ia_pts, gu, blk, 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_ia_pts"))) return rc;
    const int nm = nz * st, nwl = (nz + 1) * st;
    hipLaunchKernelGGL(k_ia_eos, dim3((nm + 255) / 256), dim3(256), 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_ia_eos"))) return rc;
    hipLaunchKernelGGL(k_ia_php, dim3((nwl + 255) / 256), dim3(256), 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_ia_php"))) return rc;
    hipLaunchKernelGGL(k_ia_cols, dim3((st + 63) / 64), dim3(64, KG), 0, sm, a);
    return lib_launched(hipGetLastError(), "k_ia_cols");
}

extern "C" int32_t ARW_API(small_step_finish_adjoint)(
    const Dom* d, const ARW_API(finish_adj)* adj, real* work, void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
    if (rc) return rc;
    if (!adj || !work) return lib_fail(ARW_E_ARG, "null argument");
    const real* lin[] = {adj->l_u, adj->l_v, adj->l_w, adj->l_thp, adj->l_php,
                         adj->l_mup};
    real* gout[] = {adj->g_u, adj->g_v, adj->g_w, adj->g_thp, adj->g_php,
                    adj->g_mup, adj->g_u_pp, adj->g_v_pp, adj->g_w_pp,
                    adj->g_th_pp, adj->g_ph_pp, adj->g_mu_pp};
    for (const real* p : lin)
        if (!p) return lib_fail(ARW_E_ARG, "null output adjoint");
    for (real* p : gout)
        if (!p) return lib_fail(ARW_E_ARG, "null input adjoint");
    const int nz = d->nz, ny = d->ny, nx = d->nx, st = ny * nx;
    if (nz < 3) return lib_fail(ARW_E_ARG, "finish adjoint needs nz >= 3");
    hipStream_t sm = (hipStream_t)stream;
    const size_t F = (size_t)(nz + 1) * (ny + 1) * (nx + 1);
    FA a;
    a.d = *d;
    for (int w = 0; w < F_COUNT; ++w) a.w[w] = work + (size_t)w * F;
    a.lu = adj->l_u; a.lv = adj->l_v; a.lw = adj->l_w; a.lthp = adj->l_thp;
    a.lphp = adj->l_php; a.lmup = adj->l_mup;
    a.gu = adj->g_u; a.gv = adj->g_v; a.gw = adj->g_w; a.gthp = adj->g_thp;
    a.gphp = adj->g_php; a.gmup = adj->g_mup;
    a.gupp = adj->g_u_pp; a.gvpp = adj->g_v_pp; a.gwpp = adj->g_w_pp;
    a.gthpp = adj->g_th_pp; a.gphpp = adj->g_ph_pp; a.gmupp = adj->g_mu_pp;
    const dim3 blk(64, 4);
    const dim3 gu((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz + 1);
    hipLaunchKernelGGL(k_fa_pts, gu, blk, 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_fa_pts"))) return rc;
    hipLaunchKernelGGL(k_fa_cols, dim3((st + 63) / 64), dim3(64, KG), 0, sm, a);
    return lib_launched(hipGetLastError(), "k_fa_cols");
}

extern "C" int32_t ARW_API(stage_adjoint_work)(int32_t nz, int32_t ny,
                                               int32_t nx, int64_t* elements)
{
    if (!elements) return lib_fail(ARW_E_ARG, "null argument");
    if (nz < 3 || ny < 1 || nx < 1) return lib_fail(ARW_E_ARG, "bad extents");
    *elements = (int64_t)S_COUNT * (nz + 1) * (ny + 1) * (nx + 1);
    return 0;
}

extern "C" int32_t ARW_API(stage_tendencies_adjoint)(
    const Dom* d, const arw_params* prm, const ARW_API(stage_adj)* adj,
    real* work, void* stream)
{
    int32_t rc = lib_check_adj_domain(d);
    if (rc) return rc;
    if (!prm || !adj || !work) return lib_fail(ARW_E_ARG, "null argument");
    if (prm->diff_6th_opt != 0)
        return lib_fail(ARW_E_UNSUPPORTED,
                        "stage adjoint: the held 6th-order tendencies are not "
                        "differentiated");
    const real* lin[] = {adj->l_ru_t, adj->l_rv_t, adj->l_rw_t, adj->l_rth_t,
                         adj->l_rph_t};
    real* gout[] = {adj->g_u, adj->g_v, adj->g_w, adj->g_thp, adj->g_php,
                    adj->g_mup};
    for (const real* p : lin)
        if (!p) return lib_fail(ARW_E_ARG, "null output adjoint");
    for (real* p : gout)
        if (!p) return lib_fail(ARW_E_ARG, "null input adjoint");
    if (d->qv && !(adj->g_qv && adj->g_qc && adj->g_qr))
        return lib_fail(ARW_E_ARG, "stage adjoint: a moist state needs g_qv, g_qc, g_qr");
    const bool diff = prm->khdif > 0.0 || prm->kvdif > 0.0;
    if (diff && (d->base3d || !d->rdzf_h || !d->rdzc_h || !d->rdzf_w
                 || !d->rdzc_w))
        return lib_fail(ARW_E_UNSUPPORTED,
                        "stage adjoint: constant-K diffusion needs a flat base "
                        "and the diffusion spacings");
    if (prm->precision_form)
        return lib_fail(ARW_E_UNSUPPORTED, "stage adjoint: plain form only");
    const int nz = d->nz, ny = d->ny, nx = d->nx, st = ny * nx;
    if (nz < 3) return lib_fail(ARW_E_ARG, "stage adjoint needs nz >= 3");
    hipStream_t sm = (hipStream_t)stream;
    const size_t F = (size_t)(nz + 1) * (ny + 1) * (nx + 1);
    SA a;
    a.d = *d;
    for (int w = 0; w < S_COUNT; ++w) a.w[w] = work + (size_t)w * F;
    a.lru = adj->l_ru_t; a.lrv = adj->l_rv_t; a.lrw = adj->l_rw_t;
    a.lrth = adj->l_rth_t; a.lrph = adj->l_rph_t;
    a.gu = adj->g_u; a.gv = adj->g_v; a.gw = adj->g_w; a.gthp = adj->g_thp;
    a.gphp = adj->g_php; a.gmup = adj->g_mup;
    a.gqv = adj->g_qv; a.gqc = adj->g_qc; a.gqr = adj->g_qr;
    const double rdx = 1.0 / prm->dx, rdy = 1.0 / prm->dy;
    a.rdx = (real)rdx; a.rdy = (real)rdy;
    a.hrdx = (real)(0.5 * rdx); a.hrdy = (real)(0.5 * rdy);
    a.qrdx = (real)(0.25 * rdx); a.qrdy = (real)(0.25 * rdy);
    a.top_lid = prm->top_lid != 0;
    if (adj->flags & ~ARW_ADJ_ACCUMULATE)
        return lib_fail(ARW_E_ARG, "unknown adjoint flags");
    a.acc = (adj->flags & ARW_ADJ_ACCUMULATE) != 0;
    // forward recompute into the work area: EOS + Omega, ru, rv, theta
    Dom& dc = a.d;
    dc.p = a.w[S_P]; dc.al = a.w[S_AL]; dc.alt = a.w[S_ALT];
    dc.ww = a.w[S_WW]; dc.dmdt_rf = a.w[S_DM]; dc.rmu_t = a.w[S_RMU];
    Sc s = make_sc(prm, prm->dt, 1);
    if ((rc = lib_launched(run_omega(dc, s, 1, sm), "k_omega (stage adjoint)")))
        return rc;
    if ((rc = lib_launched(run_faces(dc, a.w[S_RU], a.w[S_RV], sm),
                           "k_faces (stage adjoint)")))
        return rc;
    const int nm = nz * st, nwl = (nz + 1) * st;
    hipLaunchKernelGGL(k_sa_theta, dim3((nm + 255) / 256), dim3(256), 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_theta"))) return rc;
    // zeroed accumulators: the outputs and the flux gradients
    const size_t sz[] = {(size_t)nz * ny * (nx + 1), (size_t)nz * (ny + 1) * nx,
                         (size_t)nwl, (size_t)nm, (size_t)nwl, (size_t)st};
    for (int q = 0; q < 6 && !a.acc; ++q)
        if (hipMemsetAsync(gout[q], 0, sz[q] * RB, sm) != hipSuccess)
            return lib_fail(ARW_E_LAUNCH, "stage adjoint: clearing outputs");
    for (int w : {S_GRU, S_GRV, S_GWW})
        if (hipMemsetAsync(a.w[w], 0, F * RB, sm) != hipSuccess)
            return lib_fail(ARW_E_LAUNCH, "stage adjoint: clearing work");
    // adjoints of the stage fluxes handed on (the scalars' sumflux) seed
    // the flux gradients the advection transposes add to
    const real* lfl[3] = {adj->l_ru, adj->l_rv, adj->l_ww};
    const size_t nfl[3] = {(size_t)nz * ny * (nx + 1), (size_t)nz * (ny + 1) * nx,
                           (size_t)nwl};
    const int sfl[3] = {S_GRU, S_GRV, S_GWW};
    for (int q = 0; q < 3; ++q)
        if (lfl[q] && hipMemcpyAsync(a.w[sfl[q]], lfl[q], nfl[q] * RB,
                                     hipMemcpyDeviceToDevice, sm) != hipSuccess)
            return lib_fail(ARW_E_LAUNCH, "stage adjoint: flux adjoints");
    // advection of theta, u, v, w with the stage fluxes
    const real dxi = a.rdx, dyi = a.rdy;
    struct { int op; const real* q; const real* lam; real* g; const real* rdz; }
    ops[] = {{0, a.w[S_TH], a.lrth, a.gthp, d->rdnw},
             {1, d->u, a.lru, a.gu, d->rdnw},
             {2, d->v, a.lrv, a.gv, d->rdnw},
             {3, d->w, a.lrw, a.gw, d->rdn}};
    for (const auto& o : ops)
        if ((rc = lib_flux_div_adjoint(o.op, o.q, a.w[S_RU], a.w[S_RV],
                                       a.w[S_WW], o.lam, o.g, a.w[S_GRU],
                                       a.w[S_GRV], a.w[S_GWW], o.rdz, d->fnm,
                                       d->fnp, dxi, dyi, nz, ny, nx, sm)))
            return rc;
    const dim3 blk(64, 4);
    const dim3 gm((nx + 63) / 64, (ny + 3) / 4, nz);
    const dim3 gu((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz + 1);
    hipLaunchKernelGGL(k_sa_pgf, gm, blk, 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_pgf"))) return rc;
    hipLaunchKernelGGL(k_sa_pts, gu, blk, 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_pts"))) return rc;
    hipLaunchKernelGGL(k_sa_omega, dim3((st + 255) / 256), dim3(256), 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_omega"))) return rc;
    const dim3 gf((nx + 1 + 63) / 64, (ny + 1 + 3) / 4, nz);
    hipLaunchKernelGGL(k_sa_flux, gf, blk, 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_flux"))) return rc;
    hipLaunchKernelGGL(k_sa_eos, dim3((nm + 255) / 256), dim3(256), 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_eos"))) return rc;
    if (diff) {
        const dim3 gd((nx + 63) / 64, (ny + 3) / 4, nz + 1);
        hipLaunchKernelGGL(k_sa_diff, gd, blk, 0, sm, a, s);
        if ((rc = lib_launched(hipGetLastError(), "k_sa_diff"))) return rc;
    }
    hipLaunchKernelGGL(k_sa_php, dim3((nwl + 255) / 256), dim3(256), 0, sm, a);
    if ((rc = lib_launched(hipGetLastError(), "k_sa_php"))) return rc;
    hipLaunchKernelGGL(k_sa_mu, dim3((st + 255) / 256), dim3(256), 0, sm, a);
    return lib_launched(hipGetLastError(), "k_sa_mu");
}
