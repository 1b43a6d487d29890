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

__device__ __forceinline__ real phi_full(const Dom& d, int k, int c)
{
    return (d.base3d ? d.phb[k * d.ny * d.nx + c] : d.phb[k]) + d.php[k * d.ny * d.nx + c];
}

// Column arrays of the mass-column pass.  Every pass below streams the
// column once in one direction, carrying neighbour-level values in
// registers, so no level reads what another level's iteration wrote; the
// restrict qualifiers then let the loads of unrolled iterations issue
// together.
struct ColIO {
    const real* __restrict__ lth; const real* __restrict__ lp;
    const real* __restrict__ lal; const real* __restrict__ lph;
    const real* __restrict__ lw; const real* __restrict__ lmu;
    const real* __restrict__ lww;
    const real* __restrict__ T1; const real* __restrict__ AL1;
    const real* __restrict__ W1; const real* __restrict__ M1;
};

__device__ __forceinline__ void adj_col_body(
    const Dom& d, const Sc& s, const int c, const ColIO io,
    real* __restrict__ GG, real* __restrict__ LRHS, real* __restrict__ LW,
    real* __restrict__ SA, real* __restrict__ SG, real* __restrict__ LT2,
    real* __restrict__ GPH, real* __restrict__ LWW, real* __restrict__ LDVP,
    real* __restrict__ L0, real* __restrict__ gw, real* __restrict__ grw,
    real* __restrict__ gph, real* __restrict__ grph, real* __restrict__ gth,
    real* __restrict__ grth, real* __restrict__ gmu, real* __restrict__ grmu)
{
    const int nz = d.nz, st = d.ny * d.nx;
    const real half = R(0.5), one = R(1.0), g = G_F, zero = R(0.0);
    const real dt = s.dtau, eps = s.epssm;
    const real mut = d.mub2d[c] + d.mup[c];
    const real M1 = io.M1[c];
    const real K = half * dt * g * (one + eps);
    real cof = half * dt * g * (one + eps);
    cof = cof * cof;
    auto c2a_at = [&](int k) {
        return (GAMMA_F * d.p[k * st + c]) / d.alt[k * st + c];
    };
    auto chm_at = [&](int k) { return d.c1h[k] * mut + d.c2h[k]; };
    auto cfm_at = [&](int k) { return d.c1f[k] * mut + d.c2f[k]; };
    auto a_at = [&](int k) -> real {
        if (k >= 2 && k < nz)
            return -cof * d.rdn[k] * d.rdnw[k - 1] * c2a_at(k - 1)
                   / (chm_at(k - 1) * cfm_at(k - 1));
        if (k == nz && !s.top_lid)
            return R(-2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                   * c2a_at(nz - 1) / (chm_at(nz - 1) * cfm_at(nz - 1));
        return zero;
    };
    // damper factor 1/(1+dw) of level k (1 where undamped)
    real hbot = zero;
    if (s.damp) hbot = phi_full(d, nz, c) / g - s.zdamp;
    auto undamp = [&](int k) -> real {
        if (!s.damp) return one;
        const real hk = phi_full(d, k, c) / g;
        if (!(hk >= hbot)) return one;
        const real arg = R(1.5707963) * (hk - hbot) / s.zdamp;
        const real sn = (real)sin((double)arg);
        return one / (one + s.dampmag * sn * sn);
    };
    real lM1 = io.lmu[c];

    // A (up): EOS reverse, phi'' update reverse, damper, Thomas factors and
    // the forward sweep of the transposed solve
    real carry_ph = zero, gph0 = zero, vprev = zero, gprev = zero;
    auto w_in = [&](int k, real ph1) {          // lambda(w) entering the solve
        const real cfs = d.c1f[k] * (mut + M1) + d.c2f[k];
        lM1 += -K * io.W1[k * st + c] * d.c1f[k] / (cfs * cfs) * ph1;
        real lw = io.lw[k * st + c] + K / cfs * ph1;
        const real u = undamp(k);
        return s.damp ? lw * u : lw;
    };
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        const real thr = th_ref(d, k, c), alt = d.alt[x], c2a = c2a_at(k);
        const real chmp = d.c1h[k] * (mut + M1) + d.c2h[k];
        const real q = alt / (chmp * thr);
        const real X = io.T1[x] - d.c1h[k] * M1 * thr;
        const real lPP = io.lp[x];
        GG[x] = io.lth[x] + c2a * q * lPP;
        lM1 += c2a * (-q * d.c1h[k] * thr - q * X * d.c1h[k] / chmp) * lPP;
        const real lAL = io.lal[x] - c2a * lPP;
        lM1 += (-(alt * d.c1h[k]) / chmp - io.AL1[x] * d.c1h[k] / chmp) * lAL;
        const real ph1 = io.lph[x] + carry_ph + d.rdnw[k] / chmp * lAL;
        carry_ph = -d.rdnw[k] / chmp * lAL;
        if (k == 0) {
            gph0 = ph1;
            LW[c] = io.lw[c];
            continue;
        }
        LRHS[x] = ph1;
        const real lw = w_in(k, ph1);
        // Thomas factors of row k (calc_coef_w) and the transposed sweep
        const real b = one + cof * d.rdn[k] * (
            d.rdnw[k] * c2a / (chm_at(k) * cfm_at(k))
            + d.rdnw[k - 1] * c2a_at(k - 1) / (chm_at(k - 1) * cfm_at(k)));
        const real cc = -cof * d.rdn[k] * d.rdnw[k] * c2a / (chm_at(k) * cfm_at(k + 1));
        const real ak = a_at(k);
        const real al = one / (b - ak * gprev);
        const real v = lw - gprev * vprev;
        gprev = cc * al;
        vprev = v;
        LW[x] = v;
        SA[x] = al;
        SG[x] = ak;
    }
    {
        const int x = nz * st + c;
        const real ph1 = io.lph[x] + carry_ph;
        LRHS[x] = ph1;
        const real lw = w_in(nz, ph1);
        const real bn = one + R(2.0) * cof * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                        * c2a_at(nz - 1) / (chm_at(nz - 1) * cfm_at(nz));
        const real an = a_at(nz);
        LW[x] = lw - gprev * vprev;
        SA[x] = one / (bn - an * gprev);
        SG[x] = an;
    }
    // B (down): backward sweep of the transposed solve
    real carry = zero;
#pragma unroll 4
    for (int k = nz; k >= 1; --k) {
        const int x = k * st + c;
        const real v = SA[x] * (LW[x] + carry);
        LW[x] = v;
        carry = -SG[x] * v;
    }
    L0[c] = LW[c] + carry;

    // C (up): reverse of the w accumulation; la(k) = lambda of row k
    auto la_at = [&](int k) -> real {
        if (k >= 1 && k < nz) return LW[k * st + c];
        if (k == nz && !s.top_lid) return LW[nz * st + c];
        return zero;
    };
    auto lup_at = [&](int k, real la) -> real {
        if (k < 1 || k >= nz) return zero;
        return half * dt * g * d.rdn[k] * c2a_at(k) * d.rdnw[k] / chm_at(k) * la;
    };
    auto ldn_at = [&](int k, real la) -> real {
        if (k >= 1 && k < nz)
            return -(half * dt * g * d.rdn[k] * c2a_at(k - 1) * d.rdnw[k - 1]
                     / chm_at(k - 1)) * la;
        if (k == nz && !s.top_lid)
            return -(dt * g * c2a_at(nz - 1) * d.rdnw[nz - 1] * d.rdnw[nz - 1]
                     / chm_at(nz - 1)) * la;
        return zero;
    };
    real lmuave = zero;
    real la_k = zero, lup_km = zero;            // level 0: no row
    real la_kp = la_at(1);
    real ldn_k = zero;
#pragma unroll 4
    for (int k = 0; k <= nz; ++k) {
        const int x = k * st + c;
        const real la_next = (k + 2 <= nz) ? la_at(k + 2) : zero;
        const real lup_k = lup_at(k, la_k);
        const real ldn_kp = ldn_at(k + 1, la_kp);
        const real T = (lup_km - lup_k) + (ldn_k - ldn_kp);
        GPH[x] = (one - eps) * T;
        if (k >= 1) LRHS[x] = LRHS[x] + (one + eps) * T;
        const bool row = (k >= 1 && k < nz) || (k == nz && !s.top_lid);
        gw[x] = row ? la_k : zero;
        grw[x] = row ? dt * la_k : zero;
        if (k < nz) {
            real t2 = zero;
            if (k >= 1) t2 += dt * g * d.rdn[k] * c2a_at(k) * d.alt[x] * la_k;
            if (k + 1 < nz)
                t2 -= dt * g * d.rdn[k + 1] * c2a_at(k) * d.alt[x] * la_kp;
            if (k == nz - 1 && !s.top_lid)
                t2 -= dt * g * R(2.0) * d.rdnw[nz - 1] * c2a_at(nz - 1)
                      * d.alt[x] * la_kp;
            LT2[x] = t2;
        }
        if (row) lmuave += -dt * g * d.c1f[k] * la_k;
        lup_km = lup_k;
        ldn_k = ldn_kp;
        la_k = la_kp;
        la_kp = la_next;
    }

    // D (up): rhs reverse -> phi'', rph_t, w'', and Omega'' via wd
    if (s.top_lid) LRHS[nz * st + c] = zero;
    real lwd_prev = zero;
    auto lr2_at = [&](int k) -> real {
        return (k >= 1 && k <= nz) ? LRHS[k * st + c] / cfm_at(k) : zero;
    };
    real lr2_k = zero;                          // level 0
    real lr2_kp = lr2_at(1);
#pragma unroll 4
    for (int k = 0; k <= nz; ++k) {
        const int x = k * st + c;
        const real lr2_next = (k + 2 <= nz) ? lr2_at(k + 2) : zero;
        if (k == 0) {
            gph[x] = GPH[x] + gph0;
            grph[x] = zero;
        } else {
            gph[x] = GPH[x] + LRHS[x];
            grph[x] = dt * lr2_k;
            gw[x] = gw[x] + dt * half * g * (one - eps) * lr2_k;
        }
        // lambda(wd[k]) (k < nz): -dt (fnm[k] lr2[k] + fnp[k+1] lr2[k+1])
        real lwd = zero;
        if (k < nz) {
            if (k >= 1) lwd += -dt * d.fnm[k] * lr2_k;
            if (k + 1 <= nz - 1) lwd += -dt * d.fnp[k + 1] * lr2_kp;
        }
        // Omega''(k) gets cwd[k] lwd[k] + cwd[k-1] lwd[k-1]
        const real cwd = (k < nz) ? half * d.rdnw[k] * (phi_full(d, k + 1, c)
                                                        - phi_full(d, k, c))
                                  : zero;
        real ww = io.lww ? io.lww[x] : zero;
        ww += cwd * lwd + lwd_prev;
        LWW[x] = ww;
        lwd_prev = (k < nz) ? cwd * lwd : zero;
        lr2_k = lr2_kp;
        lr2_kp = lr2_next;
    }

    // E (up): theta'' adjoint (t2 term, then the theta'' update), the
    // vertical theta flux into Omega''
    real lM0old = half * (one - eps) * lmuave;
    lM1 += half * (one + eps) * lmuave;
    real gg_prev = zero;
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        const real thr = th_ref(d, k, c);
        const real chs = d.c1h[k] * (mut + M1) + d.c2h[k];
        const real t2 = half * ((one + eps) * io.T1[x] + (one - eps) * d.th_pp[x])
                        / (chs * thr);
        const real lt2 = LT2[x];
        const real lT1 = GG[x] + half * (one + eps) / (chs * thr) * lt2;
        gth[x] = half * (one - eps) / (chs * thr) * lt2 + lT1;
        grth[x] = dt * lT1;
        lM1 += -t2 * d.c1h[k] / chs * lt2;
        const real gg = -dt * lT1;
        GG[x] = gg;
        if (k >= 1) {
            const real lwdtn = d.rdnw[k - 1] * gg_prev - d.rdnw[k] * gg;
            LWW[x] = LWW[x] + (d.fnm[k] * thr + d.fnp[k] * th_ref(d, k - 1, c)) * lwdtn;
        }
        gg_prev = gg;
    }

    // F (down): Omega'' suffix sums -> mu'' and the divergence adjoint
    real acc = zero, ldmdt = zero, gr = zero;
#pragma unroll 4
    for (int k = nz - 1; k >= 0; --k) {
        real linc = zero;
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
    grmu[c] = gr;
    gmu[c] = lM1 + lM0old;
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        const int x = k * st + c;
        LDVP[x] = LDVP[x] + d.dnw[k] * ldmdt;
    }
}

// Mass-column reverse of advance_w_phi + diagnose_p and of the vertical
// parts of advance_mu_th.
__global__ __launch_bounds__(64) void k_adj_col(Dom d, Sc s, Adj a, Fw f)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ny * d.nx) return;
    ColIO io;
    io.lth = a.lth; io.lp = a.lp; io.lal = a.lal; io.lph = a.lph;
    io.lw = a.lw; io.lmu = a.lmu; io.lww = a.lww;
    io.T1 = f.T1; io.AL1 = f.AL1; io.W1 = f.W1; io.M1 = f.M1;
    adj_col_body(d, s, c, io, f.s[W_GG], f.s[W_LRHS], f.s[W_LW], f.s[W_SA],
                 f.s[W_SG], f.s[W_LT2], f.s[W_LWD], f.s[W_LWW], f.s[W_LDVP],
                 f.s[W_L0], a.gw, a.grw, a.gph, a.grph, a.gth, a.grth, a.gmu,
                 a.grmu);
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

__device__ __forceinline__ void adj_face_body(
    const Dom& d, const Sc& s, const FaceGeo& fg, const real rd,
    const real* __restrict__ lout, const real* __restrict__ GG,
    const real* __restrict__ LDVP, const real* __restrict__ L0,
    real* __restrict__ gout, real* __restrict__ grt, real* __restrict__ LB,
    real* __restrict__ LT, real* __restrict__ LS, real* __restrict__ LM,
    const FaceGeo* dup)
{
    const int nz = d.nz, st = d.ny * d.nx;
    const real half = R(0.5), dt = s.dtau, zero = R(0.0);
    const real cf[3] = {d.cf1, d.cf2, d.cf3};
    auto lam1 = [&](const FaceGeo& q, int k) -> real {
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
    const real mA = d.mub2d[fg.cA] + d.mup[fg.cA];
    const real mB = d.mub2d[fg.cB] + d.mup[fg.cB];
    const real muf = half * (mB + mA);
    real lms = zero;
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        const real l1 = lam1(fg, k);
        gout[fg.out0 + k * fg.ostride] = l1;
        grt[fg.out0 + k * fg.ostride] = dt * l1;
        if (!fg.own) continue;
        const real ldup = dup ? lam1(*dup, k) : zero;
        const real ld = -dt * l1 - dt * ldup;
        const int x = k * st + fg.slot;
        LB[x] = half * rd * (d.c1h[k] * muf + d.c2h[k]) * ld;
        real dphp = half * ((d.php[k * st + fg.cA] + d.php[(k + 1) * st + fg.cA])
                            - (d.php[k * st + fg.cB] + d.php[(k + 1) * st + fg.cB]));
        if (d.base3d)
            dphp = dphp + half * ((d.phb[k * st + fg.cA] + d.phb[(k + 1) * st + fg.cA])
                                  - (d.phb[k * st + fg.cB] + d.phb[(k + 1) * st + fg.cB]));
        const real tk = rd * dphp * ld;
        LT[x] = tk;
        lms += -half * d.c1h[k] * tk;
    }
    if (!fg.own) return;
    LM[fg.slot] = lms;
    // lambda(dpn) at full level kf: rdnw[kf-1] t[kf-1] - rdnw[kf] t[kf]
    auto ldpn = [&](int kf) -> real {
        real v = zero;
        if (kf >= 1) v += d.rdnw[kf - 1] * LT[(kf - 1) * st + fg.slot];
        if (kf <= nz - 1) v -= d.rdnw[kf] * LT[kf * st + fg.slot];
        return v;
    };
    const real l0 = ldpn(0), ltop = s.top_lid ? ldpn(nz) : zero;
    real dk = zero, dkp = ldpn(1);
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        const real dnext = (k + 2 <= nz - 1) ? ldpn(k + 2) : zero;
        real v = zero;
        if (k >= 1) v += half * d.fnm[k] * dk;
        if (k + 1 <= nz - 1) v += half * d.fnp[k + 1] * dkp;
        if (k < 3) v += half * cf[k] * l0;
        if (s.top_lid && k >= nz - 3) v += half * cf[nz - 1 - k] * ltop;
        LS[k * st + fg.slot] = v;
        dk = dkp;
        dkp = dnext;
    }
}

__global__ __launch_bounds__(128) void k_adj_faces(Dom d, Sc s, Adj a, Fw f)
{
    const int ny = d.ny, nx = d.nx, st = ny * nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = ny * (nx + 1), nv = (ny + 1) * nx;
    if (t >= nu + nv) return;
    FaceGeo fg, dg;
    bool has_dup = false;
    if (t < nu) {
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
        adj_face_body(d, s, fg, s.rdx, a.lu, f.s[W_GG], f.s[W_LDVP],
                      f.s[W_L0], a.gu, a.gru, f.s[W_LBX], f.s[W_LPH1],
                      f.s[W_LSX], f.s[W_LMX], has_dup ? &dg : nullptr);
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
        adj_face_body(d, s, fg, s.rdy, a.lv, f.s[W_GG], f.s[W_LDVP],
                      f.s[W_L0], a.gv, a.grv, f.s[W_LBY], f.s[W_LW],
                      f.s[W_LSY], f.s[W_LMY], has_dup ? &dg : nullptr);
    }
    (void)st;
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
    hipLaunchKernelGGL(k_adj_col, dim3((cols + 63) / 64), dim3(64), 0, st,
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
