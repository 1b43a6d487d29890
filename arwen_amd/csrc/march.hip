// k-marching kernels of the RK3 split-explicit step on gfx950.
//
// Every kernel assigns one thread to one vertical column (mass column,
// u face column or v face column) and marches up (or down) through the
// levels.  64-lane waves run along x, so each level of each field is one
// coalesced row; values a column needs at several levels ride in register
// windows and are read from HBM once; horizontal neighbours come from the
// neighbouring lanes' rows through L1/L2.  Expressions keep the evaluation
// order of the float32 restatement of the reference step (see common.h),
// so results are bit-identical to it.
#include "march.h"

#define R __restrict__

namespace {

__device__ __forceinline__ float mu_at(const float* R mub, const float* R mup,
                                       int c)
{
    return mub[c] + mup[c];
}

// Vertical face flux of a half-level field at w-level kf from the window
// q(kf-2), q(kf-1), q(kf), q(kf+1) (reference advection.cu:189-201).
__device__ __forceinline__ float zf(float qm2, float qm1, float q0, float qp1,
                                    float vel, int kf, int nz,
                                    const float* R fnm, const float* R fnp)
{
    if (kf == 0 || kf == nz) return 0.0f;
    if (kf == 1 || kf == nz - 1) return vel * (fnm[kf] * q0 + fnp[kf] * qm1);
    return flux3(qm2, qm1, q0, qp1, vel);
}

__device__ __forceinline__ float div3(float dfx, float dfy, float dfz,
                                      float rdx, float rdy, float rdz)
{
    float t = -dfx * rdx;
    t = t - dfy * rdy;
    return t - dfz * rdz;
}

// Large-step PGF on a face (reference dycore.cu:98-157) from values at
// level k of the two columns A (right/north) and B.
__device__ __forceinline__ float pgf_val(
    float rd, float hrd, float c1, float c2, float rdnwk, float muf, float dmu,
    float phA0, float phA1, float phB0, float phB1, float altA, float altB,
    float ppA, float ppB, float alA, float alB, float pbA, float pbB,
    int base3d, float phhA, float phhB, float dpnA0, float dpnA1, float dpnB0,
    float dpnB1)
{
    float layer = c1 * muf + c2;
    float bracket = (phA1 - phB1) + (phA0 - phB0);
    bracket = bracket + (altA + altB) * (ppA - ppB);
    if (base3d)
        bracket = bracket + (alA + alB) * (pbA - pbB);
    else
        bracket = bracket + (alA + alB) * 0.0f;
    float left = (hrd * layer) * bracket;
    float dphp = phhA - phhB;
    float hi = 0.5f * (dpnA1 + dpnB1);
    float lo = 0.5f * (dpnA0 + dpnB0);
    float vertical = rdnwk * (hi - lo) - c1 * dmu;
    return left + (rd * dphp) * vertical;
}

// Vertical window of one mass column for the large-step PGF.
struct PgfCol {
    int c;
    float pp_m, pp_0, pp_p;   // p - pb at k-1, k, k+1
    float ph_0, ph_p;         // php at k, k+1
    float gb_0, gb_p;         // phb at k, k+1 (terrain)
    float dpn_0, dpn_p;       // dpn at k, k+1
};

__device__ __forceinline__ float pp_of(const Dom& d, const float* R p,
                                       const float* R pb, int k, int c,
                                       int st)
{
    return p[k * st + c] - basev(pb, k, c, st, d.base3d);
}

__device__ __forceinline__ void pgf_init(PgfCol& X, const Dom& d, int c,
                                         int st, const float* R p,
                                         const float* R pb,
                                         const float* R php,
                                         const float* R phb)
{
    X.c = c;
    float p0 = pp_of(d, p, pb, 0, c, st), p1 = pp_of(d, p, pb, 1, c, st);
    float p2 = pp_of(d, p, pb, 2, c, st);
    X.pp_m = 0.0f;
    X.pp_0 = p0;
    X.pp_p = p1;
    X.ph_0 = php[c];
    X.gb_0 = d.base3d ? phb[c] : 0.0f;
    X.dpn_0 = (d.cf1 * p0 + d.cf2 * p1) + d.cf3 * p2;
}

// Values at level k+1 entering the window at iteration k.
__device__ __forceinline__ void pgf_load(PgfCol& X, const Dom& d, int k,
                                         int st, int top_lid,
                                         const float* R php,
                                         const float* R phb,
                                         const float* R fnm,
                                         const float* R fnp)
{
    const int nz = d.nz;
    X.ph_p = php[(k + 1) * st + X.c];
    X.gb_p = d.base3d ? phb[(k + 1) * st + X.c] : 0.0f;
    if (k + 1 == nz)
        X.dpn_p = top_lid ? d.cfn * X.pp_0 + d.cfn1 * X.pp_m : 0.0f;
    else
        X.dpn_p = fnm[k + 1] * X.pp_p + fnp[k + 1] * X.pp_0;
}

__device__ __forceinline__ void pgf_shift(PgfCol& X, const Dom& d, int k,
                                          int st, const float* R p,
                                          const float* R pb)
{
    X.pp_m = X.pp_0;
    X.pp_0 = X.pp_p;
    X.pp_p = (k + 2 < d.nz) ? pp_of(d, p, pb, k + 2, X.c, st) : 0.0f;
    X.ph_0 = X.ph_p;
    X.gb_0 = X.gb_p;
    X.dpn_0 = X.dpn_p;
}

__device__ __forceinline__ float phh_of(const PgfCol& X, int base3d)
{
    if (base3d) return 0.5f * ((X.gb_0 + X.gb_p) + (X.ph_0 + X.ph_p));
    return 0.5f * (X.ph_0 + X.ph_p);
}

// ---------------------------------------------------------------------------
// Omega and the stage EOS
// ---------------------------------------------------------------------------

__device__ __forceinline__ void eos_at(const Dom& d, int k, int c, int st,
                                       float mu, float thp_k, float php_k,
                                       float php_kp)
{
    float th = basev(d.thb, k, c, st, d.base3d) + thp_k;
    float dph = (basev(d.phb, k + 1, c, st, d.base3d) + php_kp)
                - (basev(d.phb, k, c, st, d.base3d) + php_k);
    float alt = (-dph * d.rdnw[k]) / (d.c1h[k] * mu + d.c2h[k]);
    d.alt[k * st + c] = alt;
    d.al[k * st + c] = alt - basev(d.alb, k, c, st, d.base3d);
    d.p[k * st + c] = P0_F * pow_gamma((RD_F * th) / (P0_F * alt));
}

// WRF calc_ww_cp (reference dycore.py:102-131): the column sum of the
// coupled-flux divergence, then Omega by the upward recurrence.  The sum
// also serves the acoustic substeps as the reference-momentum mass
// divergence (acoustic.cu:396-405): dnw*div(C(mu_face) u_t*) is the same
// float32 expression, so one sum serves both.  Pass 1 parks each level's
// divergence in ww; pass 2 turns it into Omega.
__global__ __launch_bounds__(256) void k_omega(Dom d, Sc s, int do_eos)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    const float* R u = d.u;
    const float* R v = d.v;
    const float* R c1h = d.c1h;
    const float* R c2h = d.c2h;
    const float* R dnw = d.dnw;
    float* R ww = d.ww;
    const float ms = mu_at(d.mub2d, d.mup, c);
    if (do_eos) {
        const float* R thp = d.thp;
        const float* R php = d.php;
        float ph_lo = php[c];
        for (int k = 0; k < nz; ++k) {
            float ph_hi = php[(k + 1) * st + c];
            eos_at(d, k, c, st, ms, thp[k * st + c], ph_lo, ph_hi);
            ph_lo = ph_hi;
        }
    }
    const float mw = mu_at(d.mub2d, d.mup, j * nx + im);
    const float me = mu_at(d.mub2d, d.mup, j * nx + ip);
    const float mS = mu_at(d.mub2d, d.mup, jm * nx + i);
    const float mN = mu_at(d.mub2d, d.mup, jp * nx + i);
    const float muxw = 0.5f * (ms + mw), muxe = 0.5f * (me + ms);
    const float muys = 0.5f * (ms + mS), muyn = 0.5f * (mN + ms);
    float dmdt = 0.0f, rf = 0.0f;
#pragma unroll 2
    for (int k = 0; k < nz; ++k) {
        float ruw = (c1h[k] * muxw + c2h[k]) * u[iU(k, j, i, ny, nx)];
        float rue = (c1h[k] * muxe + c2h[k]) * u[iU(k, j, i + 1, ny, nx)];
        float rvs = (c1h[k] * muys + c2h[k]) * v[iV(k, j, i, ny, nx)];
        float rvn = (c1h[k] * muyn + c2h[k]) * v[iV(k, j + 1, i, ny, nx)];
        float divv = dnw[k] * (s.rdx * (rue - ruw) + s.rdy * (rvn - rvs));
        dmdt = (k == 0) ? divv : dmdt + divv;
        rf = rf + divv;
        ww[(k + 1) * st + c] = divv;
    }
    ww[c] = 0.0f;
    float run = 0.0f;
    for (int k = 0; k < nz - 1; ++k) {
        float inc = (c1h[k] * dnw[k]) * dmdt + ww[(k + 1) * st + c];
        run = run + inc;
        ww[(k + 1) * st + c] = -run;
    }
    ww[nz * st + c] = 0.0f;
    d.dmdt_rf[c] = rf;
    d.rmu_t[c] = 0.0f;
}

// ---------------------------------------------------------------------------
// Slow tendencies on u and v faces (+ u''/v'' initialisation)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_tend_faces(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > nx || j > ny) return;
    const bool hu = j < ny, hv = i < nx;
    const int ic = (i >= nx) ? i - nx : i, jc = (j >= ny) ? j - ny : j;
    const int im = (ic == 0) ? nx - 1 : ic - 1;
    const int jm = (jc == 0) ? ny - 1 : jc - 1;
    const int ip = (ic + 1 == nx) ? 0 : ic + 1;
    const int jp = (jc + 1 == ny) ? 0 : jc + 1;
    int xo[7], yo[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        xo[o] = pmod(ic + o - 3, nx);
        yo[o] = pmod(jc + o - 3, ny);
    }
    const float* R u = d.u;
    const float* R v = d.v;
    const float* R ww = d.ww;
    const float* R p = d.p;
    const float* R pb = d.pb;
    const float* R php = d.php;
    const float* R phb = d.phb;
    const float* R alt = d.alt;
    const float* R al = d.al;
    const float* R c1h = d.c1h;
    const float* R c2h = d.c2h;
    const float* R fnm = d.fnm;
    const float* R fnp = d.fnp;
    const float* R rdnw = d.rdnw;
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const float* R mup0 = d.mup0;
    const int base3d = d.base3d;

    const int cA = jc * nx + ic, cW = jc * nx + im, cS = jm * nx + ic;
    const float mA = mu_at(mub, mup, cA), mW = mu_at(mub, mup, cW);
    const float mS = mu_at(mub, mup, cS);
    const float mE = mu_at(mub, mup, jc * nx + ip);
    const float mN = mu_at(mub, mup, jp * nx + ic);
    // u-face masses
    const float mux_a = 0.5f * (mW + mu_at(mub, mup, jc * nx + pmod(im - 1, nx)));
    const float mux_c = 0.5f * (mA + mW);
    const float mux_d = 0.5f * (mE + mA);
    const float muy_jW = 0.5f * (mW + mu_at(mub, mup, jm * nx + im));
    const float muy_jA = 0.5f * (mA + mS);
    const float muy_pW = 0.5f * (mu_at(mub, mup, jp * nx + im) + mW);
    const float muy_pA = 0.5f * (mN + mA);
    // v-face masses
    const float muy_a = 0.5f * (mS + mu_at(mub, mup, pmod(jm - 1, ny) * nx + ic));
    const float muy_c = 0.5f * (mA + mS);
    const float muy_d = 0.5f * (mN + mA);
    const float mux_jS = 0.5f * (mS + mu_at(mub, mup, jm * nx + im));
    const float mux_jA = 0.5f * (mA + mW);
    const float mux_pS = 0.5f * (mu_at(mub, mup, jm * nx + ip) + mS);
    const float mux_pA = 0.5f * (mE + mA);
    // PGF and init masses
    const float dmu_u = 0.5f * (mup[cA] + mup[cW]);
    const float dmu_v = 0.5f * (mup[cA] + mup[cS]);
    const float mtA = mub[cA] + mup0[cA];
    const float mtf_u = 0.5f * (mtA + (mub[cW] + mup0[cW]));
    const float mtf_v = 0.5f * (mtA + (mub[cS] + mup0[cS]));
    const float msf_u = 0.5f * (mA + mW), msf_v = 0.5f * (mA + mS);

    PgfCol A, W, S;
    pgf_init(A, d, cA, st, p, pb, php, phb);
    if (hu) pgf_init(W, d, cW, st, p, pb, php, phb);
    if (hv) pgf_init(S, d, cS, st, p, pb, php, phb);

    float uz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float vz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int kk = 0; kk < 3 && kk < nz; ++kk) {
        if (hu) uz[2 + kk] = u[iU(kk, jc, ic, ny, nx)];
        if (hv) vz[2 + kk] = v[iV(kk, jc, ic, ny, nx)];
    }
    float wA0 = ww[cA], wW0 = hu ? ww[cW] : 0.0f, wS0 = hv ? ww[cS] : 0.0f;

    for (int k = 0; k < nz; ++k) {
        const float c1 = c1h[k], c2 = c2h[k];
        pgf_load(A, d, k, st, s.top_lid, php, phb, fnm, fnp);
        const float wA1 = ww[(k + 1) * st + cA];
        const float altA = alt[k * st + cA], alA = al[k * st + cA];
        const float pbA = base3d ? pb[k * st + cA] : 0.0f;

        if (hu) {
            pgf_load(W, d, k, st, s.top_lid, php, phb, fnm, fnp);
            const float wW1 = ww[(k + 1) * st + cW];
            float x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) x[o] = u[iU(k, jc, xo[o], ny, nx)];
#pragma unroll
            for (int o = 0; o < 7; ++o)
                y[o] = (o == 3) ? x[3] : u[iU(k, yo[o], ic, ny, nx)];
            const float u_b = (ic == 0) ? u[iU(k, jc, nx, ny, nx)] : x[3];
            const float u_d = (ic + 1 == nx) ? u[iU(k, jc, nx, ny, nx)] : x[4];
            float ru_a = (c1 * mux_a + c2) * x[2];
            float ru_b = (c1 * mux_c + c2) * u_b;
            float ru_c = (c1 * mux_c + c2) * x[3];
            float ru_d = (c1 * mux_d + c2) * u_d;
            float vl = 0.5f * (ru_a + ru_b), vr = 0.5f * (ru_c + ru_d);
            float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vl);
            float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], vr);
            float rv_jW = (c1 * muy_jW + c2) * v[iV(k, jc, im, ny, nx)];
            float rv_jA = (c1 * muy_jA + c2) * v[iV(k, jc, ic, ny, nx)];
            float rv_pW = (c1 * muy_pW + c2) * v[iV(k, jc + 1, im, ny, nx)];
            float rv_pA = (c1 * muy_pA + c2) * v[iV(k, jc + 1, ic, ny, nx)];
            float v0 = 0.5f * (rv_jW + rv_jA), v1 = 0.5f * (rv_pW + rv_pA);
            float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
            float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
            float wz0 = 0.5f * (wW0 + wA0), wz1 = 0.5f * (wW1 + wA1);
            float fz0 = zf(uz[0], uz[1], uz[2], uz[3], wz0, k, nz, fnm, fnp);
            float fz1 = zf(uz[1], uz[2], uz[3], uz[4], wz1, k + 1, nz, fnm, fnp);
            float t = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy,
                           rdnw[k]);
            const float altW = alt[k * st + cW], alW = al[k * st + cW];
            const float pbW = base3d ? pb[k * st + cW] : 0.0f;
            t = t - pgf_val(s.rdx, s.hrdx, c1, c2, rdnw[k], mux_c, dmu_u,
                            A.ph_0, A.ph_p, W.ph_0, W.ph_p, altA, altW,
                            A.pp_0, W.pp_0, alA, alW, pbA, pbW, base3d,
                            phh_of(A, base3d), phh_of(W, base3d), A.dpn_0,
                            A.dpn_p, W.dpn_0, W.dpn_p);
            if (s.diff) {
                float fc = x[3];
                float lap = s.kh * (s.dxi2 * ((x[4] - 2.0f * fc) + x[2])
                                    + s.dyi2 * ((y[4] - 2.0f * fc) + y[2]));
                float fup = (k < nz - 1) ? (uz[3] - fc) * d.rdzf_h[k] : 0.0f;
                float fdn = (k > 0) ? (fc - uz[1]) * d.rdzf_h[k - 1] : 0.0f;
                lap = lap + (s.kv * (fup - fdn)) * d.rdzc_h[k];
                t = t + (c1 * msf_u + c2) * lap;
            }
            const int ix = iU(k, jc, i, ny, nx);
            d.ru_t[ix] = t;
            const float uo = (i == ic) ? x[3] : u[ix];
            d.u_pp[ix] = (c1 * mtf_u + c2) * d.u0[ix] - (c1 * msf_u + c2) * uo;
            wW0 = wW1;
            pgf_shift(W, d, k, st, p, pb);
            uz[0] = uz[1]; uz[1] = uz[2]; uz[2] = uz[3]; uz[3] = uz[4];
            uz[4] = (k + 3 < nz) ? u[iU(k + 3, jc, ic, ny, nx)] : 0.0f;
        }
        if (hv) {
            pgf_load(S, d, k, st, s.top_lid, php, phb, fnm, fnp);
            const float wS1 = ww[(k + 1) * st + cS];
            float x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) y[o] = v[iV(k, yo[o], ic, ny, nx)];
#pragma unroll
            for (int o = 0; o < 7; ++o)
                x[o] = (o == 3) ? y[3] : v[iV(k, jc, xo[o], ny, nx)];
            float ru_jS = (c1 * mux_jS + c2) * u[iU(k, jm, ic, ny, nx)];
            float ru_jA = (c1 * mux_jA + c2) * u[iU(k, jc, ic, ny, nx)];
            float ru_pS = (c1 * mux_pS + c2) * u[iU(k, jm, ic + 1, ny, nx)];
            float ru_pA = (c1 * mux_pA + c2) * u[iU(k, jc, ic + 1, ny, nx)];
            float u0c = 0.5f * (ru_jS + ru_jA), u1c = 0.5f * (ru_pS + ru_pA);
            float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0c);
            float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1c);
            const float v_b = (jc == 0) ? v[iV(k, ny, ic, ny, nx)] : y[3];
            const float v_d = (jc + 1 == ny) ? v[iV(k, ny, ic, ny, nx)] : y[4];
            float rv_a = (c1 * muy_a + c2) * y[2];
            float rv_b = (c1 * muy_c + c2) * v_b;
            float rv_c = (c1 * muy_c + c2) * y[3];
            float rv_d = (c1 * muy_d + c2) * v_d;
            float vl = 0.5f * (rv_a + rv_b), vr = 0.5f * (rv_c + rv_d);
            float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], vl);
            float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], vr);
            float wz0 = 0.5f * (wS0 + wA0), wz1 = 0.5f * (wS1 + wA1);
            float fz0 = zf(vz[0], vz[1], vz[2], vz[3], wz0, k, nz, fnm, fnp);
            float fz1 = zf(vz[1], vz[2], vz[3], vz[4], wz1, k + 1, nz, fnm, fnp);
            float t = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy,
                           rdnw[k]);
            const float altS = alt[k * st + cS], alS = al[k * st + cS];
            const float pbS = base3d ? pb[k * st + cS] : 0.0f;
            t = t - pgf_val(s.rdy, s.hrdy, c1, c2, rdnw[k], muy_c, dmu_v,
                            A.ph_0, A.ph_p, S.ph_0, S.ph_p, altA, altS,
                            A.pp_0, S.pp_0, alA, alS, pbA, pbS, base3d,
                            phh_of(A, base3d), phh_of(S, base3d), A.dpn_0,
                            A.dpn_p, S.dpn_0, S.dpn_p);
            if (s.diff) {
                float fc = y[3];
                float lap = s.kh * (s.dxi2 * ((x[4] - 2.0f * fc) + x[2])
                                    + s.dyi2 * ((y[4] - 2.0f * fc) + y[2]));
                float fup = (k < nz - 1) ? (vz[3] - fc) * d.rdzf_h[k] : 0.0f;
                float fdn = (k > 0) ? (fc - vz[1]) * d.rdzf_h[k - 1] : 0.0f;
                lap = lap + (s.kv * (fup - fdn)) * d.rdzc_h[k];
                t = t + (c1 * msf_v + c2) * lap;
            }
            const int ix = iV(k, j, ic, ny, nx);
            d.rv_t[ix] = t;
            const float vo = (j == jc) ? y[3] : v[ix];
            d.v_pp[ix] = (c1 * mtf_v + c2) * d.v0[ix] - (c1 * msf_v + c2) * vo;
            wS0 = wS1;
            pgf_shift(S, d, k, st, p, pb);
            vz[0] = vz[1]; vz[1] = vz[2]; vz[2] = vz[3]; vz[3] = vz[4];
            vz[4] = (k + 3 < nz) ? v[iV(k + 3, jc, ic, ny, nx)] : 0.0f;
        }
        wA0 = wA1;
        pgf_shift(A, d, k, st, p, pb);
    }
}

// ---------------------------------------------------------------------------
// Slow tendencies in the mass column (theta, w, phi) + column init
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_tend_cols(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const int c = j * nx + i;
    const int im = (i == 0) ? nx - 1 : i - 1, ip = (i + 1 == nx) ? 0 : i + 1;
    const int jm = (j == 0) ? ny - 1 : j - 1, jp = (j + 1 == ny) ? 0 : j + 1;
    int xo[7], yo[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        xo[o] = j * nx + pmod(i + o - 3, nx);
        yo[o] = pmod(j + o - 3, ny) * nx + i;
    }
    const float* R u = d.u;
    const float* R v = d.v;
    const float* R w = d.w;
    const float* R thp = d.thp;
    const float* R thb = d.thb;
    const float* R php = d.php;
    const float* R phb = d.phb;
    const float* R ww = d.ww;
    const float* R p = d.p;
    const float* R pb = d.pb;
    const float* R c1h = d.c1h;
    const float* R c2h = d.c2h;
    const float* R c1f = d.c1f;
    const float* R c2f = d.c2f;
    const float* R fnm = d.fnm;
    const float* R fnp = d.fnp;
    const float* R rdn = d.rdn;
    const float* R rdnw = d.rdnw;
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const int b3 = d.base3d;

    const float mA = mu_at(mub, mup, c), mW = mu_at(mub, mup, j * nx + im);
    const float mE = mu_at(mub, mup, j * nx + ip);
    const float mS = mu_at(mub, mup, jm * nx + i);
    const float mN = mu_at(mub, mup, jp * nx + i);
    const float mux_i = 0.5f * (mA + mW), mux_ip = 0.5f * (mE + mA);
    const float muy_j = 0.5f * (mA + mS), muy_jp = 0.5f * (mN + mA);
    const float mupc = mup[c];
    const float mt = mub[c] + d.mup0[c];
    const float mupp = d.mup0[c] - mupc;
    d.mu_pp[c] = mupp;
    const float qx = s.qrdx, qy = s.qrdy;

    auto th_at = [&](int k, int cc) {
        return basev(thb, k, cc, st, b3) + thp[k * st + cc];
    };
    auto ph_at = [&](int k, int cc) {
        return basev(phb, k, cc, st, b3) + php[k * st + cc];
    };

    // windows
    float thz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};   // theta k-2..k+2
    float wz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};    // w k-2..k+2
    for (int kk = 0; kk < 3; ++kk) {
        if (kk < nz) thz[2 + kk] = th_at(kk, c);
        wz[2 + kk] = w[kk * st + c];
    }
    float ww_m = 0.0f, ww_0 = ww[c], ww_p = ww[st + c];
    float ph_m = 0.0f, ph_0 = ph_at(0, c), ph_p = ph_at(1, c);
    float pp_m = 0.0f;
    float ru_i_m = 0.0f, ru_ip_m = 0.0f, rv_j_m = 0.0f, rv_jp_m = 0.0f;
    float u_i_m = 0.0f, u_i_mm = 0.0f, u_ip_m = 0.0f, u_ip_mm = 0.0f;
    float v_j_m = 0.0f, v_j_mm = 0.0f, v_jp_m = 0.0f, v_jp_mm = 0.0f;
    float thp_m = 0.0f, thp_0 = thp[c], thp_p = (nz > 1) ? thp[st + c] : 0.0f;
    float phpp_0 = d.php0[c] - php[c];      // ph'' init at level k

    for (int k = 0; k <= nz; ++k) {
        const bool mass = k < nz;
        float pp_0 = mass ? p[k * st + c] - basev(pb, k, c, st, b3) : 0.0f;
        float ru_i = 0.0f, ru_ip = 0.0f, rv_j = 0.0f, rv_jp = 0.0f;
        float u_i = 0.0f, u_ip = 0.0f, v_j = 0.0f, v_jp = 0.0f;
        if (mass) {
            ru_i = (c1h[k] * mux_i + c2h[k]) * u[iU(k, j, i, ny, nx)];
            ru_ip = (c1h[k] * mux_ip + c2h[k]) * u[iU(k, j, i + 1, ny, nx)];
            rv_j = (c1h[k] * muy_j + c2h[k]) * v[iV(k, j, i, ny, nx)];
            rv_jp = (c1h[k] * muy_jp + c2h[k]) * v[iV(k, j + 1, i, ny, nx)];
            u_i = u[iU(k, j, i, ny, nx)];
            u_ip = u[iU(k, j, ip, ny, nx)];
            v_j = v[iV(k, j, i, ny, nx)];
            v_jp = v[iV(k, jp, i, ny, nx)];
        }

        // --- rw_t: advection + buoyancy + diffusion ---
        float t = 0.0f;
        float wx2 = 0.0f, wx4 = 0.0f, wy2 = 0.0f, wy4 = 0.0f;
        if (k >= 1 && k < nz) {
            float x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) {
                x[o] = (o == 3) ? wz[2] : w[k * st + xo[o]];
                y[o] = (o == 3) ? wz[2] : w[k * st + yo[o]];
            }
            wx2 = x[2]; wx4 = x[4]; wy2 = y[2]; wy4 = y[4];
            float u0v = fnm[k] * ru_i + fnp[k] * ru_i_m;
            float u1v = fnm[k] * ru_ip + fnp[k] * ru_ip_m;
            float v0v = fnm[k] * rv_j + fnp[k] * rv_j_m;
            float v1v = fnm[k] * rv_jp + fnp[k] * rv_jp_m;
            float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0v);
            float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1v);
            float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0v);
            float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1v);
            float fz[2];
            const float wwv[3] = {ww_m, ww_0, ww_p};
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                int m = k - 1 + q;
                float vel = 0.5f * (wwv[q] + wwv[q + 1]);
                if (m == 0 || m == nz - 1)
                    fz[q] = 0.5f * vel * (wz[q + 2] + wz[q + 1]);
                else
                    fz[q] = flux3(wz[q], wz[q + 1], wz[q + 2], wz[q + 3], vel);
            }
            t = div3(fx1 - fx0, fy1 - fy0, fz[1] - fz[0], s.rdx, s.rdy, rdn[k]);
            float term = rdn[k] * (pp_0 - pp_m) - c1f[k] * mupc;
            t = t + G_F * term;
        } else if (k == nz) {
            float pressure = ((1.0f * 2.0f) * rdnw[nz - 1]) * (0.0f - pp_m);
            float term = pressure - c1f[nz] * mupc;
            t = t + G_F * term;
        }
        if (s.diff) {
            float lap = 0.0f;
            if (k >= 1 && k < nz) {
                float fc = wz[2];
                lap = s.kh * (s.dxi2 * ((wx4 - 2.0f * fc) + wx2)
                              + s.dyi2 * ((wy4 - 2.0f * fc) + wy2));
                float fup = (wz[3] - fc) * d.rdzf_w[k];
                float fdn = (fc - wz[1]) * d.rdzf_w[k - 1];
                lap = lap + (s.kv * (fup - fdn)) * d.rdzc_w[k];
            }
            t = t + (c1f[k] * mA + c2f[k]) * lap;
        }
        d.rw_t[k * st + c] = t;

        // --- rph_t: WRF rhs_ph ---
        float r = 0.0f;
        if (k >= 1) {
            const bool top = (k == nz);
            if (!top) {
                float wd_lo = ((0.5f * (ww_0 + ww_m)) * rdnw[k - 1]) * (ph_0 - ph_m);
                float wd_hi = ((0.5f * (ww_p + ww_0)) * rdnw[k]) * (ph_p - ph_0);
                float omega = fnm[k] * wd_hi + fnp[k] * wd_lo;
                r = r - omega;
            }
            float mass_k = c1f[k] * mA + c2f[k];
            r = r + mass_k * (G_F * wz[2]);
            float qxk = top ? 2.0f * qx : qx, qyk = top ? 2.0f * qy : qy;
            float wi, wip, wj, wjp;
            if (top) {
                wi = d.cfn * u_i_m + d.cfn1 * u_i_mm;
                wip = d.cfn * u_ip_m + d.cfn1 * u_ip_mm;
                wj = d.cfn * v_j_m + d.cfn1 * v_j_mm;
                wjp = d.cfn * v_jp_m + d.cfn1 * v_jp_mm;
            } else {
                wi = u_i + u_i_m;
                wip = u_ip + u_ip_m;
                wj = v_j + v_j_m;
                wjp = v_jp + v_jp_m;
            }
            float fcx_i = (c1f[k] * mux_i + c2f[k]) * wi;
            float fcx_ip = (c1f[k] * mux_ip + c2f[k]) * wip;
            float fcy_j = (c1f[k] * muy_j + c2f[k]) * wj;
            float fcy_jp = (c1f[k] * muy_jp + c2f[k]) * wjp;
            float p0 = ph_0;
            float fx = fcx_i * (p0 - ph_at(k, j * nx + im));
            float fxp = fcx_ip * (ph_at(k, j * nx + ip) - p0);
            r = r - qxk * (fxp + fx);
            float fy = fcy_j * (p0 - ph_at(k, jm * nx + i));
            float fyp = fcy_jp * (ph_at(k, jp * nx + i) - p0);
            r = r - qyk * (fyp + fy);
        }
        d.rph_t[k * st + c] = r;

        // --- w'' and phi'' initialisation at w level k ---
        d.w_pp[k * st + c] = (c1f[k] * mt + c2f[k]) * d.w0[k * st + c]
                             - (c1f[k] * mA + c2f[k]) * wz[2];
        d.ph_pp[k * st + c] = phpp_0;

        // --- mass level k: rth_t and the theta''/p''/alpha'' init ---
        float phpp_p = 0.0f;
        if (mass) {
            float qxv[7], qyv[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) {
                qxv[o] = (o == 3) ? thz[2] : th_at(k, xo[o]);
                qyv[o] = (o == 3) ? thz[2] : th_at(k, yo[o]);
            }
            float fx0 = flux5(qxv[0], qxv[1], qxv[2], qxv[3], qxv[4], qxv[5], ru_i);
            float fx1 = flux5(qxv[1], qxv[2], qxv[3], qxv[4], qxv[5], qxv[6], ru_ip);
            float fy0 = flux5(qyv[0], qyv[1], qyv[2], qyv[3], qyv[4], qyv[5], rv_j);
            float fy1 = flux5(qyv[1], qyv[2], qyv[3], qyv[4], qyv[5], qyv[6], rv_jp);
            float fz0 = zf(thz[0], thz[1], thz[2], thz[3], ww_0, k, nz, fnm, fnp);
            float fz1 = zf(thz[1], thz[2], thz[3], thz[4], ww_p, k + 1, nz, fnm, fnp);
            float t2 = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy, rdnw[k]);
            if (s.diff) {
                float fc = thp_0;
                float lap = s.kh * (s.dxi2 * ((thp[k * st + j * nx + ip] - 2.0f * fc)
                                              + thp[k * st + j * nx + im])
                                    + s.dyi2 * ((thp[k * st + jp * nx + i] - 2.0f * fc)
                                                + thp[k * st + jm * nx + i]));
                float fup = (k < nz - 1) ? (thp_p - fc) * d.rdzf_h[k] : 0.0f;
                float fdn = (k > 0) ? (fc - thp_m) * d.rdzf_h[k - 1] : 0.0f;
                lap = lap + (s.kv * (fup - fdn)) * d.rdzc_h[k];
                t2 = t2 + (c1h[k] * mA + c2h[k]) * lap;
            }
            d.rth_t[k * st + c] = t2;

            const int h = k * st + c;
            phpp_p = d.php0[h + st] - php[h + st];
            float cht = c1h[k] * mt + c2h[k];
            float chs = c1h[k] * mA + c2h[k];
            float thbk = basev(thb, k, c, st, b3);
            float tht = thbk + d.thp0[h];
            float ths = thbk + thp_0;
            float thpp = cht * tht - chs * ths;
            d.th_pp[h] = thpp;
            float altk = d.alt[h];
            float c2a = (GAMMA_F * p[h]) / altk;
            float al = -((altk * (c1h[k] * mupp)) + rdnw[k] * (phpp_p - phpp_0)) / cht;
            if (b3) d.al_pp[h] = al;
            float p5 = (altk * (thpp - (c1h[k] * mupp) * ths)) / (cht * ths);
            d.p_pp[h] = c2a * (p5 - al);
        }

        // shift windows
        phpp_0 = phpp_p;
        thz[0] = thz[1]; thz[1] = thz[2]; thz[2] = thz[3]; thz[3] = thz[4];
        thz[4] = (k + 3 < nz) ? th_at(k + 3, c) : 0.0f;
        wz[0] = wz[1]; wz[1] = wz[2]; wz[2] = wz[3]; wz[3] = wz[4];
        wz[4] = (k + 3 <= nz) ? w[(k + 3) * st + c] : 0.0f;
        ww_m = ww_0; ww_0 = ww_p;
        ww_p = (k + 2 <= nz) ? ww[(k + 2) * st + c] : 0.0f;
        ph_m = ph_0; ph_0 = ph_p;
        ph_p = (k + 2 <= nz) ? ph_at(k + 2, c) : 0.0f;
        pp_m = pp_0;
        ru_i_m = ru_i; ru_ip_m = ru_ip; rv_j_m = rv_j; rv_jp_m = rv_jp;
        u_i_mm = u_i_m; u_i_m = u_i; u_ip_mm = u_ip_m; u_ip_m = u_ip;
        v_j_mm = v_j_m; v_j_m = v_j; v_jp_mm = v_jp_m; v_jp_m = v_jp;
        thp_m = thp_0; thp_0 = thp_p;
        thp_p = (k + 2 < nz) ? thp[(k + 2) * st + c] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Acoustic substep, part 1: u''/v'' forward step on faces
// ---------------------------------------------------------------------------

// Damped-pressure window of one column for the acoustic PGF.
struct AcCol {
    int c;
    float pd_0, pd_p;          // damped p'' at k, k+1 (own column)
    float ph_0, ph_p;          // ph'' at k, k+1
    float hp_0, hp_p;          // php at k, k+1
    float gb_0, gb_p;          // phb at k, k+1 (terrain)
};

__device__ __forceinline__ float pd_of(const float* R pp, const float* R po,
                                       int ix, float smdiv, int use_pold)
{
    float a = pp[ix];
    float b = use_pold ? po[ix] : a;
    return a + smdiv * (a - b);
}

__global__ __launch_bounds__(256) void k_uv(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > nx || j > ny) return;
    const bool hu = j < ny, hv = i < nx;
    const int ic = (i >= nx) ? i - nx : i, jc = (j >= ny) ? j - ny : j;
    const int im = (ic == 0) ? nx - 1 : ic - 1;
    const int jm = (jc == 0) ? ny - 1 : jc - 1;
    const float* R p_pp = d.p_pp;
    const float* R p_old = d.p_pp_old;
    const float* R ph_pp = d.ph_pp;
    const float* R php = d.php;
    const float* R phb = d.phb;
    const float* R alt = d.alt;
    const float* R al_pp = d.al_pp;
    const float* R pb = d.pb;
    const float* R c1h = d.c1h;
    const float* R c2h = d.c2h;
    const float* R fnm = d.fnm;
    const float* R fnp = d.fnp;
    const float* R rdnw = d.rdnw;
    const float smdiv = s.smdiv;
    const int usep = s.use_pold;
    const int b3 = d.base3d;
    const float dtau = s.dtau;

    const int cA = jc * nx + ic, cW = jc * nx + im, cS = jm * nx + ic;
    const float mA = mu_at(d.mub2d, d.mup, cA);
    const float muf_u = 0.5f * (mu_at(d.mub2d, d.mup, cW) + mA);
    const float muf_v = 0.5f * (mu_at(d.mub2d, d.mup, cS) + mA);
    const float mppA = d.mu_pp[cA];
    const float dmu_u = d.mu_pp[cW] + mppA, dmu_v = d.mu_pp[cS] + mppA;

    // per-column vertical data: damped p'' sums per face need levels
    // k-2..k+1 (surface and lid extrapolations)
    auto pdA = [&](int k) { return pd_of(p_pp, p_old, k * st + cA, smdiv, usep); };
    auto pdW = [&](int k) { return pd_of(p_pp, p_old, k * st + cW, smdiv, usep); };
    auto pdS = [&](int k) { return pd_of(p_pp, p_old, k * st + cS, smdiv, usep); };
    // S windows: S(k-2), S(k-1), S(k), S(k+1)
    float pa[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pw[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float ps[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    pa[2] = pdA(0); pa[3] = pdA(1);
    if (hu) { pw[2] = pdW(0); pw[3] = pdW(1); }
    if (hv) { ps[2] = pdS(0); ps[3] = pdS(1); }
    const float pa2 = pdA(2);
    const float pw2 = hu ? pdW(2) : 0.0f, ps2 = hv ? pdS(2) : 0.0f;
    // dpn at level 0
    float dnu_0 = 0.0f, dnv_0 = 0.0f;
    if (hu) {
        float S0 = pa[2] + pw[2], S1 = pa[3] + pw[3], S2 = pa2 + pw2;
        dnu_0 = 0.5f * ((d.cf1 * S0 + d.cf2 * S1) + d.cf3 * S2);
    }
    if (hv) {
        float S0 = pa[2] + ps[2], S1 = pa[3] + ps[3], S2 = pa2 + ps2;
        dnv_0 = 0.5f * ((d.cf1 * S0 + d.cf2 * S1) + d.cf3 * S2);
    }
    float phA0 = ph_pp[cA], phW0 = hu ? ph_pp[cW] : 0.0f, phS0 = hv ? ph_pp[cS] : 0.0f;
    float hpA0 = php[cA], hpW0 = hu ? php[cW] : 0.0f, hpS0 = hv ? php[cS] : 0.0f;
    float gbA0 = b3 ? phb[cA] : 0.0f, gbW0 = (b3 && hu) ? phb[cW] : 0.0f;
    float gbS0 = (b3 && hv) ? phb[cS] : 0.0f;

    for (int k = 0; k < nz; ++k) {
        const float c1 = c1h[k], c2 = c2h[k];
        const int a = k * st + cA;
        const float phA1 = ph_pp[a + st], hpA1 = php[a + st];
        const float gbA1 = b3 ? phb[a + st] : 0.0f;
        const float altA = alt[a];
        if (hu) {
            const int b = k * st + cW;
            const float phW1 = ph_pp[b + st], hpW1 = php[b + st];
            const float gbW1 = b3 ? phb[b + st] : 0.0f;
            float Sk = pa[2] + pw[2];
            float dn_up;
            if (k + 1 == nz) {
                if (s.top_lid) {
                    float Sm1 = pa[1] + pw[1], Sm2 = pa[0] + pw[0];
                    dn_up = 0.5f * ((d.cf1 * Sk + d.cf2 * Sm1) + d.cf3 * Sm2);
                } else {
                    dn_up = 0.0f;
                }
            } else {
                float Sp = pa[3] + pw[3];
                dn_up = 0.5f * (fnm[k + 1] * Sp + fnp[k + 1] * Sk);
            }
            float dph = (phA1 - phW1) + (phA0 - phW0);
            float dpe = pa[2] - pw[2];
            float brk = dph + (altA + alt[b]) * dpe;
            if (b3) brk = brk + (al_pp[a] + al_pp[b]) * (pb[a] - pb[b]);
            float dpxy = ((0.5f * s.rdx) * (c1 * muf_u + c2)) * brk;
            float phsA = hpA0 + hpA1, phsB = hpW0 + hpW1;
            float dphp = 0.5f * (phsA - phsB);
            if (b3) dphp = dphp + 0.5f * ((gbA0 + gbA1) - (gbW0 + gbW1));
            dpxy = dpxy + (s.rdx * dphp) * (rdnw[k] * (dn_up - dnu_0)
                                            - (0.5f * c1) * dmu_u);
            const int ix = iU(k, jc, i, ny, nx);
            d.u_pp[ix] = d.u_pp[ix] + dtau * (d.ru_t[ix] - dpxy);
            dnu_0 = dn_up;
            phW0 = phW1; hpW0 = hpW1; gbW0 = gbW1;
        }
        if (hv) {
            const int b = k * st + cS;
            const float phS1 = ph_pp[b + st], hpS1 = php[b + st];
            const float gbS1 = b3 ? phb[b + st] : 0.0f;
            float Sk = pa[2] + ps[2];
            float dn_up;
            if (k + 1 == nz) {
                if (s.top_lid) {
                    float Sm1 = pa[1] + ps[1], Sm2 = pa[0] + ps[0];
                    dn_up = 0.5f * ((d.cf1 * Sk + d.cf2 * Sm1) + d.cf3 * Sm2);
                } else {
                    dn_up = 0.0f;
                }
            } else {
                float Sp = pa[3] + ps[3];
                dn_up = 0.5f * (fnm[k + 1] * Sp + fnp[k + 1] * Sk);
            }
            float dph = (phA1 - phS1) + (phA0 - phS0);
            float dpe = pa[2] - ps[2];
            float brk = dph + (altA + alt[b]) * dpe;
            if (b3) brk = brk + (al_pp[a] + al_pp[b]) * (pb[a] - pb[b]);
            float dpxy = ((0.5f * s.rdy) * (c1 * muf_v + c2)) * brk;
            float phsA = hpA0 + hpA1, phsB = hpS0 + hpS1;
            float dphp = 0.5f * (phsA - phsB);
            if (b3) dphp = dphp + 0.5f * ((gbA0 + gbA1) - (gbS0 + gbS1));
            dpxy = dpxy + (s.rdy * dphp) * (rdnw[k] * (dn_up - dnv_0)
                                            - (0.5f * c1) * dmu_v);
            const int ix = iV(k, j, ic, ny, nx);
            d.v_pp[ix] = d.v_pp[ix] + dtau * (d.rv_t[ix] - dpxy);
            dnv_0 = dn_up;
            phS0 = phS1; hpS0 = hpS1; gbS0 = gbS1;
        }
        phA0 = phA1; hpA0 = hpA1; gbA0 = gbA1;
        // advance the damped-pressure windows to k+1
        pa[0] = pa[1]; pa[1] = pa[2]; pa[2] = pa[3];
        pa[3] = (k + 2 < nz) ? pdA(k + 2) : 0.0f;
        if (hu) {
            pw[0] = pw[1]; pw[1] = pw[2]; pw[2] = pw[3];
            pw[3] = (k + 2 < nz) ? pdW(k + 2) : 0.0f;
        }
        if (hv) {
            ps[0] = ps[1]; ps[1] = ps[2]; ps[2] = ps[3];
            ps[3] = (k + 2 < nz) ? pdS(k + 2) : 0.0f;
        }
    }
}

// ---------------------------------------------------------------------------
// Acoustic substep, part 2: one thread per mass column
// ---------------------------------------------------------------------------

// mu''/Omega''/(mu theta)'' forward step (reference acoustic.cu:329-456),
// the implicit w''/phi'' solve with the kinematic lower boundary, the
// damp_opt=3 damper and the fused linearised EOS (acoustic.cu:774-982).
// The tridiagonal LU factors (calc_coef_w, acoustic.cu:675-759) are formed
// on the way up from p/alt of the stage reference state -- the same
// float32 expressions the reference evaluates once per stage -- so no
// coefficient arrays stream in; gamma is parked for the way down.  Between
// the sweeps w'' holds the forward-eliminated values and ph'' the phi''
// right-hand side.  pout receives the new p''.
__global__ __launch_bounds__(256) void k_column(Dom d, Sc s, float* R pout,
                                               int write_al, int write_ww)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    const int b3 = d.base3d;
    const float eps = s.epssm, dtau = s.dtau;
    const float one_p = 1.0f + eps, one_m = 1.0f - eps;
    const float* R u_pp = d.u_pp;
    const float* R v_pp = d.v_pp;
    const float* R thp = d.thp;
    const float* R thb = d.thb;
    const float* R php = d.php;
    const float* R phb = d.phb;
    const float* R alt = d.alt;
    const float* R p = d.p;
    const float* R rth_t = d.rth_t;
    const float* R rph_t = d.rph_t;
    const float* R rw_t = d.rw_t;
    const float* R c1h = d.c1h;
    const float* R c2h = d.c2h;
    const float* R c1f = d.c1f;
    const float* R c2f = d.c2f;
    const float* R dnw = d.dnw;
    const float* R rdnw = d.rdnw;
    const float* R rdn = d.rdn;
    const float* R fnm = d.fnm;
    const float* R fnp = d.fnp;
    float* R th_pp = d.th_pp;
    float* R w_pp = d.w_pp;
    float* R ph_pp = d.ph_pp;
    float* R gam = d.gam;

    auto dvp_at = [&](int k) {
        return s.rdx * (u_pp[iU(k, j, i + 1, ny, nx)] - u_pp[iU(k, j, i, ny, nx)])
               + s.rdy * (v_pp[iV(k, j + 1, i, ny, nx)] - v_pp[iV(k, j, i, ny, nx)]);
    };
    const float mut = mu_at(d.mub2d, d.mup, c);
    const float mupp_old = d.mu_pp[c];
    float dmdt_pp = 0.0f;
#pragma unroll 2
    for (int k = 0; k < nz; ++k) dmdt_pp = dmdt_pp + dnw[k] * dvp_at(k);
    const float rmu = d.rmu_t[c];
    const float mupp = mupp_old + dtau * ((dmdt_pp + d.dmdt_rf[c]) + rmu);
    d.mu_pp[c] = mupp;
    const float muts = mut + mupp;
    const float muave = 0.5f * (one_p * mupp + one_m * mupp_old);
    float cof = 0.5f * dtau * G_F * one_p;
    cof = cof * cof;

    auto thr = [&](int kk, int cc) {
        return basev(thb, kk, cc, st, b3) + thp[kk * st + cc];
    };
    auto phs = [&](int kk) { return basev(phb, kk, c, st, b3) + php[kk * st + c]; };
    auto cfm_at = [&](int kk) { return c1f[kk] * mut + c2f[kk]; };

    // kinematic lower boundary on w'' from the updated u''/v''
    float w0;
    {
        auto lw = [&](const float* R a, int ix0, int ix1, int ix2) {
            return (d.cf1 * a[ix0] + d.cf2 * a[ix1]) + d.cf3 * a[ix2];
        };
        float ue = lw(u_pp, iU(0, j, i + 1, ny, nx), iU(1, j, i + 1, ny, nx),
                      iU(2, j, i + 1, ny, nx));
        float uw = lw(u_pp, iU(0, j, i, ny, nx), iU(1, j, i, ny, nx),
                      iU(2, j, i, ny, nx));
        float vn = lw(v_pp, iV(0, j + 1, i, ny, nx), iV(1, j + 1, i, ny, nx),
                      iV(2, j + 1, i, ny, nx));
        float vs = lw(v_pp, iV(0, j, i, ny, nx), iV(1, j, i, ny, nx),
                      iV(2, j, i, ny, nx));
        const float* R ht = d.ht;
        float h = ht[c];
        w0 = (0.5f * s.rdy) * ((ht[jp * nx + i] - h) * vn + (h - ht[jm * nx + i]) * vs)
             + (0.5f * s.rdx) * ((ht[j * nx + ip] - h) * ue + (h - ht[j * nx + im]) * uw);
        w_pp[c] = w0;
    }

    // upward sweep
    if (write_ww) d.ww_pp[c] = 0.0f;
    float ww_k = 0.0f, wwk = 0.0f, wdtn_k = 0.0f;
    float th_m = thr(0, c), ph_m = phs(0), wd_prev = 0.0f;
    float rhs_a = 0.0f, rhs_b = 0.0f;
    float pho_a = 0.0f, pho_b = ph_pp[c];
    float t2_a = 0.0f, t2_b = 0.0f;
    float c2a_a = 0.0f, c2a_b = 0.0f;
    float alt_a = 0.0f, alt_b = 0.0f;
    float chm_a = 0.0f, chm_b = 0.0f;
    float wold_b = 0.0f;
    float wf_prev = w0;
    float gam_prev = 0.0f;
    const float ghalf_eps = (0.5f * G_F) * one_m;
    const float c_force = (0.5f * dtau) * G_F;
    const float c_buoy = dtau * G_F;

    auto force_row = [&](int r, float rhs_dn, float rhs_c, float rhs_up,
                         float ph_dn, float ph_c, float ph_up, float wold,
                         float t2dn, float t2c, float c2dn, float c2c,
                         float altdn, float altc, float chmdn, float chmc) {
        float dph_up = one_p * (rhs_up - rhs_c) + one_m * (ph_up - ph_c);
        float dph_dn = one_p * (rhs_c - rhs_dn) + one_m * (ph_c - ph_dn);
        float t1 = dtau * rw_t[r * st + c]
                   + (c_force * rdn[r])
                         * ((c2c * rdnw[r]) / chmc * dph_up
                            - (c2dn * rdnw[r - 1]) / chmdn * dph_dn);
        float t3 = c_buoy * (rdn[r] * ((c2c * altc) * t2c - (c2dn * altdn) * t2dn)
                             - c1f[r] * muave);
        return wold + (t1 + t3);
    };
    // LU factors of row r (1 <= r <= nz-1) from the windowed c2a/chm
    auto lu_row = [&](int r, float c2dn, float c2c, float chmdn, float chmc,
                      float& alpha_r) {
        float a_r = 0.0f;
        if (r >= 2)
            a_r = -cof * rdn[r] * rdnw[r - 1] * c2dn / (chmdn * cfm_at(r - 1));
        float b = 1.0f + cof * rdn[r] * (rdnw[r] * c2c / (chmc * cfm_at(r))
                                         + rdnw[r - 1] * c2dn / (chmdn * cfm_at(r)));
        float cc = -cof * rdn[r] * rdnw[r] * c2c / (chmc * cfm_at(r + 1));
        alpha_r = 1.0f / (b - a_r * gam_prev);
        gam_prev = cc * alpha_r;
        gam[r * st + c] = gam_prev;
        return a_r;
    };

    for (int m = 0; m < nz; ++m) {
        const int h = m * st + c;
        float dvp = dvp_at(m);
        float ww_kp1;
        if (m + 1 < nz) {
            wwk = wwk - dnw[m] * (c1h[m] * (dmdt_pp + rmu) + dvp);
            ww_kp1 = wwk;
        } else {
            ww_kp1 = 0.0f;
        }
        if (write_ww) d.ww_pp[h + st] = ww_kp1;
        float th_up = (m + 1 < nz) ? thr(m + 1, c) : 0.0f;
        float wdtn_kp1 = (m + 1 < nz)
            ? ww_kp1 * (fnm[m + 1] * th_up + fnp[m + 1] * th_m) : 0.0f;
        float t_c = th_m;
        float hx = (0.5f * s.rdx) * (u_pp[iU(m, j, i + 1, ny, nx)] * (thr(m, j * nx + ip) + t_c)
                                     - u_pp[iU(m, j, i, ny, nx)] * (t_c + thr(m, j * nx + im)));
        float hy = (0.5f * s.rdy) * (v_pp[iV(m, j + 1, i, ny, nx)] * (thr(m, jp * nx + i) + t_c)
                                     - v_pp[iV(m, j, i, ny, nx)] * (t_c + thr(m, jm * nx + i)));
        float tho = th_pp[h];
        float thn = tho + dtau * (rth_t[h] - ((hx + hy) + rdnw[m] * (wdtn_kp1 - wdtn_k)));
        th_pp[h] = thn;
        float chs = c1h[m] * muts + c2h[m];
        float t2 = (0.5f * (one_p * thn + one_m * tho)) / (chs * t_c);
        float ph_up = phs(m + 1);
        float wd = ((0.5f * (ww_kp1 + ww_k)) * rdnw[m]) * (ph_up - ph_m);
        float altm = alt[h];
        float c2a = GAMMA_F * p[h] / altm;
        float chm = c1h[m] * mut + c2h[m];
        float rhs_c = 0.0f, pho_c = pho_b, wold_c = 0.0f;
        if (m >= 1) {
            pho_c = ph_pp[h];
            wold_c = w_pp[h];
            float r = dtau * (rph_t[h] + ghalf_eps * wold_c);
            r = r - dtau * (fnm[m] * wd + fnp[m] * wd_prev);
            rhs_c = pho_c + r / cfm_at(m);
            ph_pp[h] = rhs_c;
        }
        if (m >= 2) {
            const int r = m - 1;
            float alpha_r;
            float a_r = lu_row(r, c2a_a, c2a_b, chm_a, chm_b, alpha_r);
            float wr = force_row(r, rhs_a, rhs_b, rhs_c, pho_a, pho_b, pho_c,
                                 wold_b, t2_a, t2_b, c2a_a, c2a_b, alt_a,
                                 alt_b, chm_a, chm_b);
            float wf = (wr - a_r * wf_prev) * alpha_r;
            w_pp[r * st + c] = wf;
            wf_prev = wf;
        }
        rhs_a = rhs_b; rhs_b = rhs_c;
        pho_a = pho_b; pho_b = pho_c;
        t2_a = t2_b; t2_b = t2;
        c2a_a = c2a_b; c2a_b = c2a;
        alt_a = alt_b; alt_b = altm;
        chm_a = chm_b; chm_b = chm;
        wold_b = wold_c;
        ww_k = ww_kp1;
        wdtn_k = wdtn_kp1;
        th_m = th_up;
        ph_m = ph_up;
        wd_prev = wd;
    }
    {
        const int f = nz * st + c;
        const float pho_c = ph_pp[f];
        const float wold_c = w_pp[f];
        float r = dtau * (rph_t[f] + ghalf_eps * wold_c);
        float rhs_c = pho_c + r / cfm_at(nz);
        if (s.top_lid) rhs_c = 0.0f;
        ph_pp[f] = rhs_c;
        const int rr = nz - 1;
        float alpha_r;
        float a_r = lu_row(rr, c2a_a, c2a_b, chm_a, chm_b, alpha_r);
        float wr = force_row(rr, rhs_a, rhs_b, rhs_c, pho_a, pho_b, pho_c,
                             wold_b, t2_a, t2_b, c2a_a, c2a_b, alt_a, alt_b,
                             chm_a, chm_b);
        float wfr = (wr - a_r * wf_prev) * alpha_r;
        w_pp[rr * st + c] = wfr;
        float a_top = 0.0f;
        float wtop;
        if (s.top_lid) {
            wtop = 0.0f;
        } else {
            a_top = -2.0f * cof * rdnw[nz - 1] * rdnw[nz - 1] * c2a_b
                    / (chm_b * cfm_at(nz - 1));
            float dph_dn = one_p * (rhs_c - rhs_b) + one_m * (pho_c - pho_b);
            float bterm = dtau * G_F * c2a_b * rdnw[nz - 1] * rdnw[nz - 1]
                          / chm_b * dph_dn;
            float cterm = dtau * G_F * (2.0f * rdnw[nz - 1] * c2a_b * alt_b * t2_b
                                        + c1f[nz] * muave);
            wtop = wold_c + ((dtau * rw_t[f] - bterm) - cterm);
        }
        float b_top = 1.0f + 2.0f * cof * rdnw[nz - 1] * rdnw[nz - 1] * c2a_b
                             / (chm_b * cfm_at(nz));
        float alpha_top = 1.0f / (b_top - a_top * gam_prev);
        w_pp[f] = (wtop - a_top * wfr) * alpha_top;
    }

    // downward sweep: back substitution, damper, phi'' update, EOS
    float htop = 0.0f, hbot = 0.0f;
    if (s.damp) {
        htop = phs(nz) / G_F;
        hbot = htop - s.zdamp;
    }
    const float c_ph = ((0.5f * dtau) * G_F) * one_p;
    float wb_next = 0.0f, phn_up = 0.0f;
    for (int k = nz; k >= 0; --k) {
        const int f = k * st + c;
        float phn;
        if (k >= 1) {
            float wb = w_pp[f];
            if (k < nz) wb = wb - gam[f] * wb_next;
            wb_next = wb;
            float wfin = wb;
            if (s.damp) {
                float hk = phs(k) / G_F;
                if (hk >= hbot) {
                    float arg = (1.5707963f * (hk - hbot)) / s.zdamp;
                    float sn = (float)sin((double)arg);
                    float dw = (s.dampmag * sn) * sn;
                    wfin = (wb - (dw * cfm_at(k)) * d.w[f]) / (1.0f + dw);
                }
            }
            w_pp[f] = wfin;
            float cfs = c1f[k] * muts + c2f[k];
            phn = ph_pp[f] + (c_ph * wfin) / cfs;
            ph_pp[f] = phn;
        } else {
            phn = ph_pp[f];
        }
        if (k < nz) {
            float chm2 = c1h[k] * muts + c2h[k];
            float th_ref = thr(k, c);
            float altk = alt[f];
            float c2a = GAMMA_F * p[f] / altk;
            float al = -((altk * (c1h[k] * mupp)) + rdnw[k] * (phn_up - phn)) / chm2;
            if (write_al) d.al_pp[f] = al;
            pout[f] = c2a * ((altk * (th_pp[f] - (c1h[k] * mupp) * th_ref))
                             / (chm2 * th_ref) - al);
        }
        phn_up = phn;
    }
}

// ---------------------------------------------------------------------------
// Small-step finish (reference dycore.cu:1087-1198) + EOS of the new state
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_finish(Dom d, FinishDst o)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > nx || j > ny) return;
    const bool hu = j < ny, hv = i < nx, hc = hu && hv;
    const int ic = (i >= nx) ? i - nx : i, jc = (j >= ny) ? j - ny : j;
    const int im = (ic == 0) ? nx - 1 : ic - 1;
    const int jm = (jc == 0) ? ny - 1 : jc - 1;
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const float* R mu_pp = d.mu_pp;
    const float* R c1h = d.c1h;
    const float* R c2h = d.c2h;
    const float* R c1f = d.c1f;
    const float* R c2f = d.c2f;
    const int b3 = d.base3d;
    const int cA = jc * nx + ic;
    const float msA = mub[cA] + mup[cA], mnA = msA + mu_pp[cA];
    float msf_u = 0.0f, mnf_u = 0.0f, msf_v = 0.0f, mnf_v = 0.0f;
    if (hu) {
        const int cB = jc * nx + im;
        const float msB = mub[cB] + mup[cB], mnB = msB + mu_pp[cB];
        msf_u = 0.5f * (msA + msB);
        mnf_u = 0.5f * (mnA + mnB);
    }
    if (hv) {
        const int cB = jm * nx + ic;
        const float msB = mub[cB] + mup[cB], mnB = msB + mu_pp[cB];
        msf_v = 0.5f * (msA + msB);
        mnf_v = 0.5f * (mnA + mnB);
    }
    const float mup_new = hc ? mup[cA] + mu_pp[cA] : 0.0f;
    const float mu_eos = hc ? mub[cA] + mup_new : 0.0f;
    float php_lo = 0.0f, thp_lo = 0.0f;
    for (int k = 0; k <= nz; ++k) {
        if (k < nz) {
            const float c1 = c1h[k], c2 = c2h[k];
            if (hu) {
                const int ix = iU(k, jc, i, ny, nx);
                o.u[ix] = ((c1 * msf_u + c2) * d.u[ix] + d.u_pp[ix])
                          / (c1 * mnf_u + c2);
            }
            if (hv) {
                const int ix = iV(k, j, ic, ny, nx);
                o.v[ix] = ((c1 * msf_v + c2) * d.v[ix] + d.v_pp[ix])
                          / (c1 * mnf_v + c2);
            }
        }
        if (hc) {
            const int f = k * st + cA;
            o.w[f] = ((c1f[k] * msA + c2f[k]) * d.w[f] + d.w_pp[f])
                     / (c1f[k] * mnA + c2f[k]);
            const float php_new = d.php[f] + d.ph_pp[f];
            o.php[f] = php_new;
            if (k >= 1) {
                const int h = f - st;
                float th = basev(d.thb, k - 1, cA, st, b3) + thp_lo;
                float dph = (basev(d.phb, k, cA, st, b3) + php_new)
                            - (basev(d.phb, k - 1, cA, st, b3) + php_lo);
                float alt = (-dph * d.rdnw[k - 1]) / (c1h[k - 1] * mu_eos + c2h[k - 1]);
                d.alt[h] = alt;
                d.al[h] = alt - basev(d.alb, k - 1, cA, st, b3);
                d.p[h] = P0_F * pow_gamma((RD_F * th) / (P0_F * alt));
            }
            if (k < nz) {
                float thbk = basev(d.thb, k, cA, st, b3);
                float num = (c1h[k] * msA + c2h[k]) * (thbk + d.thp[f]) + d.th_pp[f];
                float thn = num / (c1h[k] * mnA + c2h[k]) - thbk;
                o.thp[f] = thn;
                thp_lo = thn;
            }
            php_lo = php_new;
        }
    }
    if (hc) o.mup[cA] = mup_new;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline dim3 cols(const Dom& d) { return dim3((d.ny * d.nx + 255) / 256); }
static inline dim3 faces(const Dom& d)
{
    return dim3((d.nx + 1 + 63) / 64, (d.ny + 1 + 3) / 4);
}
static inline dim3 masses(const Dom& d)
{
    return dim3((d.nx + 63) / 64, (d.ny + 3) / 4);
}

hipError_t run_omega(const Dom& d, const Sc& s, int do_eos, hipStream_t st)
{
    hipLaunchKernelGGL(k_omega, cols(d), dim3(256), 0, st, d, s, do_eos);
    return hipGetLastError();
}

hipError_t run_tend_faces(const Dom& d, const Sc& s, hipStream_t st)
{
    hipLaunchKernelGGL(k_tend_faces, faces(d), dim3(64, 4), 0, st, d, s);
    return hipGetLastError();
}

hipError_t run_tend_cols(const Dom& d, const Sc& s, hipStream_t st)
{
    hipLaunchKernelGGL(k_tend_cols, masses(d), dim3(64, 4), 0, st, d, s);
    return hipGetLastError();
}

hipError_t run_uv(const Dom& d, const Sc& s, hipStream_t st)
{
    hipLaunchKernelGGL(k_uv, faces(d), dim3(64, 4), 0, st, d, s);
    return hipGetLastError();
}

hipError_t run_column(const Dom& d, const Sc& s, float* pout, int write_al,
                      int write_ww, hipStream_t st)
{
    hipLaunchKernelGGL(k_column, cols(d), dim3(256), 0, st, d, s, pout,
                       write_al, write_ww);
    return hipGetLastError();
}

hipError_t run_finish(const Dom& d, const FinishDst& o, hipStream_t st)
{
    hipLaunchKernelGGL(k_finish, faces(d), dim3(64, 4), 0, st, d, o);
    return hipGetLastError();
}
