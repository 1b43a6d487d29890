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
#include <type_traits>

#include "march.h"

#define R __restrict__

namespace {

// Compile-time level kinds: loops peel their special levels so the
// interior level body has no branches.
template <int V>
using Kind = std::integral_constant<int, V>;
enum { LV_FIRST = 0, LV_SECOND = 1, LV_THIRD = 2, LV_MID = 3, LV_LAST = 4 };

__device__ __forceinline__ float mu_at(const float* R mub, const float* R mup,
                                       int c)
{
    return mub[c] + mup[c];
}

// Vertical face flux of a half-level field at w-level kf from the window
// q(kf-2), q(kf-1), q(kf), q(kf+1) (reference advection.cu:189-201).
template <class CP>
__device__ __forceinline__ float zf(float qm2, float qm1, float q0, float qp1,
                                    float vel, int kf, int nz, CP fnm, CP fnp)
{
    // both forms are evaluated and one is selected, so a level loop built
    // on zf stays one basic block (fnm/fnp have nz entries)
    const int kc = kf < nz ? kf : nz - 1;
    const float lin = vel * (fnm[kc] * q0 + fnp[kc] * qm1);
    float cub = flux3(qm2, qm1, q0, qp1, vel);
    keep(cub);
    const float r = (kf == 1 || kf == nz - 1) ? lin : cub;
    return (kf == 0 || kf == nz) ? 0.0f : r;
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

// ---------------------------------------------------------------------------
// Omega and the stage EOS
// ---------------------------------------------------------------------------

// WRF calc_ww_cp (reference dycore.py:102-131): the column sum of the
// coupled-flux divergence, then Omega by the upward recurrence.  The sum
// also serves the acoustic substeps as the reference-momentum mass
// divergence (acoustic.cu:396-405): dnw*div(C(mu_face) u_t*) is the same
// float32 expression, so one sum serves both.  Pass 1 parks each level's
// divergence in ww; pass 2 turns it into Omega.  EOS: the stage-start
// calc_p_alpha (diagnostics.cu:37-89) when the stage state is new.
template <bool EOS, bool B3>
__global__ __launch_bounds__(256) void k_omega(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    kfloat* dnw = kview(d.dnw);
    kfloat* rdnw = kview(d.rdnw);
    unsigned bc = 4u * c, bu = 4u * (j * (nx + 1) + i), bu1 = bu + 4u;
    unsigned bvn = 4u * (c + nx);
    const float ms = d.mub2d[c] + d.mup[c];
    if (EOS) {
        kfloat* thb1 = kview(d.thb);
        kfloat* phb1 = kview(d.phb);
        kfloat* alb1 = kview(d.alb);
        float ph_lo = gld(d.php, 0, bc);
        float gb_lo = B3 ? gld(d.phb, 0, bc) : phb1[0];
#pragma unroll 2
        for (int k = 0; k < nz; ++k) {
            colb_opaque(bc);
            const unsigned h = k * st;
            const float ph_hi = gld(d.php, h + st, bc);
            const float gb_hi = B3 ? gld(d.phb, h + st, bc) : phb1[k + 1];
            float th = (B3 ? gld(d.thb, h, bc) : thb1[k]) + gld(d.thp, h, bc);
            if (d.qv) th = th * (1.0f + RVOVRD_F * gld(d.qv, h, bc));
            const float albk = B3 ? gld(d.alb, h, bc) : alb1[k];
            float dph = (gb_hi + ph_hi) - (gb_lo + ph_lo);
            float alt = (-dph * rdnw[k]) / (c1h[k] * ms + c2h[k]);
            gst(d.alt, h, bc, alt);
            gst(d.al, h, bc, alt - albk);
            gst(d.p, h, bc, P0_F * pow_gamma((RD_F * th) / (P0_F * alt)));
            ph_lo = ph_hi;
            gb_lo = gb_hi;
        }
    }
    const float mw = d.mub2d[j * nx + im] + d.mup[j * nx + im];
    const float me = d.mub2d[j * nx + ip] + d.mup[j * nx + ip];
    const float mS = d.mub2d[jm * nx + i] + d.mup[jm * nx + i];
    const float mN = d.mub2d[jp * nx + i] + d.mup[jp * nx + i];
    const float muxw = 0.5f * (ms + mw), muxe = 0.5f * (me + ms);
    const float muys = 0.5f * (ms + mS), muyn = 0.5f * (mN + ms);
    float dmdt = 0.0f, rf = 0.0f;
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        colb_opaque(bc); colb_opaque(bu); colb_opaque(bu1); colb_opaque(bvn);
        const float c1 = c1h[k], c2 = c2h[k];
        float ruw = (c1 * muxw + c2) * gld(d.u, k * stu, bu);
        float rue = (c1 * muxe + c2) * gld(d.u, k * stu, bu1);
        float rvs = (c1 * muys + c2) * gld(d.v, k * stv, bc);
        float rvn = (c1 * muyn + c2) * gld(d.v, k * stv, bvn);
        float divv = dnw[k] * (s.rdx * (rue - ruw) + s.rdy * (rvn - rvs));
        dmdt = (k == 0) ? divv : dmdt + divv;
        rf = rf + divv;
        gst(d.ww, (k + 1) * st, bc, divv);
    }
    gst(d.ww, 0, bc, 0.0f);
    float run = 0.0f;
#pragma unroll 4
    for (int k = 0; k < nz - 1; ++k) {
        colb_opaque(bc);
        const unsigned h = (k + 1) * st;
        float inc = (c1h[k] * dnw[k]) * dmdt + gld(d.ww, h, bc);
        run = run + inc;
        gst(d.ww, h, bc, -run);
    }
    gst(d.ww, nz * st, bc, 0.0f);
    d.dmdt_rf[c] = rf;
    d.rmu_t[c] = 0.0f;
}

// ---------------------------------------------------------------------------
// Row-segment waves.  The stencil kernels run 64x4 workgroups in which
// each 64-lane wave covers 64 consecutive columns of ONE row, so the row
// index is wave-uniform: every row offset (the y stencil rows j-3..j+3,
// j+-1, the periodic wrap in y) is scalar arithmetic, and each lane keeps
// only its column offsets (the x stencil i-3..i+3 with periodic wrap).
// ---------------------------------------------------------------------------
struct RowSeg {
    int i, j;          // column (per lane), row (wave-uniform)
    bool live;
};

// Row segment of workgroup (blockIdx.x - first) of a grid whose first
// `first` workgroups (a multiple of ARW_NXCD) do other work.
__device__ __forceinline__ RowSeg row_seg_from(int nx, int ny, int first)
{
    const int nbx = (nx + 63) / 64;
    const int n = (int)gridDim.x - first, b = (int)blockIdx.x - first;
    const int L = (b % ARW_NXCD) * (n / ARW_NXCD) + b / ARW_NXCD;
    const int by = L / nbx, bx = L - by * nbx;
    RowSeg r;
    r.i = bx * 64 + (int)threadIdx.x;
    r.j = __builtin_amdgcn_readfirstlane(by * 4 + (int)threadIdx.y);
    r.live = r.i < nx && r.j < ny;
    return r;
}

__device__ __forceinline__ RowSeg row_seg(int nx, int ny)
{
    const int nbx = (nx + 63) / 64;
    const int L = xcd_block();
    const int by = L / nbx, bx = L - by * nbx;
    RowSeg r;
    r.i = bx * 64 + (int)threadIdx.x;
    r.j = __builtin_amdgcn_readfirstlane(by * 4 + (int)threadIdx.y);
    r.live = r.i < nx && r.j < ny;
    return r;
}

static inline dim3 rowseg_grid(int nx, int ny)
{
    return dim3(xcd_grid((unsigned)((nx + 63) / 64) * ((ny + 3) / 4)));
}

// Vertical window of one mass column for the large-step PGF: p - pb at
// k-1, k, k+1; php, phb and dpn at w levels k, k+1.
struct PgfWin {
    float pp_m = 0.0f, pp_0 = 0.0f, pp_p = 0.0f;
    float ph_0 = 0.0f, ph_p = 0.0f, gb_0 = 0.0f, gb_p = 0.0f;
    float dpn_0 = 0.0f, dpn_p = 0.0f;
};

// Slow tendencies on the u face (j, i) and v face (j, i): 5th/3rd-order
// advection (reference advection.cu:424-794), large-step PGF
// (dycore.cu:98-157), constant-K diffusion (diffusion.cu:22-51), and the
// u''/v'' small-step initialisation (dycore.cu:947-1010).  The periodic
// duplicate faces are finished by k_tend_faces_edge.
template <bool B3, bool DIFF>
__global__ __launch_bounds__(256) void k_tend_faces(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int pu = nx + 1;                       // u row pitch
    const int stu = ny * pu, stv = (ny + 1) * nx;
    const RowSeg rs = row_seg(nx, ny);
    if (!rs.live) return;
    const int i = rs.i, j = rs.j;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    kfloat* fnm = kview(d.fnm);
    kfloat* fnp = kview(d.fnp);
    kfloat* rdnw = kview(d.rdnw);
    kfloat* rdzf_h = kview(d.rdzf_h);
    kfloat* rdzc_h = kview(d.rdzc_h);
    // wave-uniform rows
    int yo[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) yo[o] = pmod(j + o - 3, ny);
    const int jm = yo[2], jp = yo[4];
    const int jb = (j == 0) ? ny : j;            // v_b row (duplicate at j=0)
    // per-lane column byte offsets
    unsigned cx[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) cx[o] = 4u * pmod(i + o - 3, nx);
    unsigned ci1 = 4u * (i + 1);                 // unwrapped east face
    unsigned cub = 4u * ((i == 0) ? nx : i);     // u_b column
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(cx[o]);
        colb_opaque(ci1); colb_opaque(cub);
    };
    unsigned& cC = cx[3];                        // centre, west
    unsigned& cW = cx[2];
    // mass columns: A = (j, i), W = (j, i-1), S = (j-1, i)
    const int rA = j * nx, rS = jm * nx;
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const float* R mup0 = d.mup0;
    auto mu2 = [&](int row, unsigned cb) {
        return gld(mub, row, cb) + gld(mup, row, cb);
    };
    const float mA = mu2(rA, cC), mW = mu2(rA, cW), mS = mu2(rS, cC);
    const float mE = mu2(rA, cx[4]), mN = mu2(jp * nx, cC);
    // u-face masses
    const float mux_a = 0.5f * (mW + mu2(rA, cx[1]));
    const float mux_c = 0.5f * (mA + mW);
    const float mux_d = 0.5f * (mE + mA);
    const float muy_jW = 0.5f * (mW + mu2(rS, cW));
    const float muy_jA = 0.5f * (mA + mS);
    const float muy_pW = 0.5f * (mu2(jp * nx, cW) + mW);
    const float muy_pA = 0.5f * (mN + mA);
    // v-face masses
    const float muy_a = 0.5f * (mS + mu2(yo[1] * nx, cC));
    const float muy_c = 0.5f * (mA + mS);
    const float muy_d = 0.5f * (mN + mA);
    const float mux_jS = 0.5f * (mS + mu2(rS, cW));
    const float mux_jA = 0.5f * (mA + mW);
    const float mux_pS = 0.5f * (mu2(rS, cx[4]) + mS);
    const float mux_pA = 0.5f * (mE + mA);
    // PGF and init masses
    const float dmu_u = 0.5f * (gld(mup, rA, cC) + gld(mup, rA, cW));
    const float dmu_v = 0.5f * (gld(mup, rA, cC) + gld(mup, rS, cC));
    const float mtA = gld(mub, rA, cC) + gld(mup0, rA, cC);
    const float mtf_u = 0.5f * (mtA + (gld(mub, rA, cW) + gld(mup0, rA, cW)));
    const float mtf_v = 0.5f * (mtA + (gld(mub, rS, cC) + gld(mup0, rS, cC)));
    const float msf_u = 0.5f * (mA + mW), msf_v = 0.5f * (mA + mS);

    // PGF windows of A, W, S
    auto pp_at = [&](int k, int row, unsigned cb) {
        const int l = k * st + row;
        return gld(d.p, l, cb) - (B3 ? gld(d.pb, l, cb) : kview(d.pb)[k]);
    };
    auto win_init = [&](PgfWin& X, int row, unsigned cb) {
        float p0 = pp_at(0, row, cb), p1 = pp_at(1, row, cb);
        float p2 = pp_at(2, row, cb);
        X.pp_0 = p0;
        X.pp_p = p1;
        X.ph_0 = gld(d.php, row, cb);
        X.gb_0 = B3 ? gld(d.phb, row, cb) : 0.0f;
        X.dpn_0 = (d.cf1 * p0 + d.cf2 * p1) + d.cf3 * p2;
    };
    auto win_load = [&](PgfWin& X, int k, int row, unsigned cb, bool top) {
        X.ph_p = gld(d.php, (k + 1) * st + row, cb);
        X.gb_p = B3 ? gld(d.phb, (k + 1) * st + row, cb) : 0.0f;
        const int kp = top ? nz - 1 : k + 1;
        const float lid = s.top_lid ? d.cfn * X.pp_0 + d.cfn1 * X.pp_m : 0.0f;
        const float mid = fnm[kp] * X.pp_p + fnp[kp] * X.pp_0;
        X.dpn_p = top ? lid : mid;
    };
    auto win_shift = [&](PgfWin& X, int k, int row, unsigned cb) {
        X.pp_m = X.pp_0;
        X.pp_0 = X.pp_p;
        const float np = pp_at(k + 2 < nz ? k + 2 : nz - 1, row, cb);
        X.pp_p = (k + 2 < nz) ? np : 0.0f;
        X.ph_0 = X.ph_p;
        X.gb_0 = X.gb_p;
        X.dpn_0 = X.dpn_p;
    };
    auto phh = [&](const PgfWin& X) {
        if (B3) return 0.5f * ((X.gb_0 + X.gb_p) + (X.ph_0 + X.ph_p));
        return 0.5f * (X.ph_0 + X.ph_p);
    };
    PgfWin A, W, S;
    win_init(A, rA, cC);
    win_init(W, rA, cW);
    win_init(S, rS, cC);

    float uz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float vz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int kk = 0; kk < 3 && kk < nz; ++kk) {
        uz[2 + kk] = gld(d.u, kk * stu + j * pu, cC);
        vz[2 + kk] = gld(d.v, kk * stv + j * nx, cC);
    }
    float wA0 = gld(d.ww, rA, cC), wW0 = gld(d.ww, rA, cW);
    float wS0 = gld(d.ww, rS, cC);

    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const bool top = k + 1 == nz;
        const float c1 = c1h[k], c2 = c2h[k];
        const int lm = k * st;
        win_load(A, k, rA, cC, top);
        win_load(W, k, rA, cW, top);
        win_load(S, k, rS, cC, top);
        const float wA1 = gld(d.ww, lm + st + rA, cC);
        const float wW1 = gld(d.ww, lm + st + rA, cW);
        const float wS1 = gld(d.ww, lm + st + rS, cC);
        const float altA = gld(d.alt, lm + rA, cC), alA = gld(d.al, lm + rA, cC);
        const float altW = gld(d.alt, lm + rA, cW), alW = gld(d.al, lm + rA, cW);
        const float altS = gld(d.alt, lm + rS, cC), alS = gld(d.al, lm + rS, cC);
        const float pbA = B3 ? gld(d.pb, lm + rA, cC) : 0.0f;
        const float pbW = B3 ? gld(d.pb, lm + rA, cW) : 0.0f;
        const float pbS = B3 ? gld(d.pb, lm + rS, cC) : 0.0f;
        const int lu = k * stu, lv = k * stv;
        // ---- u face ----
        {
            float x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) x[o] = gld(d.u, lu + j * pu, cx[o]);
#pragma unroll
            for (int o = 0; o < 7; ++o)
                y[o] = (o == 3) ? x[3] : gld(d.u, lu + yo[o] * pu, cC);
            const float u_b = gld(d.u, lu + j * pu, cub);
            const float u_d = gld(d.u, lu + j * pu, ci1);
            float ru_a = (c1 * mux_a + c2) * x[2];
            float ru_b = (c1 * mux_c + c2) * u_b;
            float ru_c = (c1 * mux_c + c2) * x[3];
            float ru_d = (c1 * mux_d + c2) * u_d;
            float vl = 0.5f * (ru_a + ru_b), vr = 0.5f * (ru_c + ru_d);
            float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vl);
            float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], vr);
            float rv_jW = (c1 * muy_jW + c2) * gld(d.v, lv + j * nx, cW);
            float rv_jA = (c1 * muy_jA + c2) * gld(d.v, lv + j * nx, cC);
            float rv_pW = (c1 * muy_pW + c2) * gld(d.v, lv + (j + 1) * nx, cW);
            float rv_pA = (c1 * muy_pA + c2) * gld(d.v, lv + (j + 1) * nx, cC);
            float v0 = 0.5f * (rv_jW + rv_jA), v1 = 0.5f * (rv_pW + rv_pA);
            float fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
            float fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
            float wz0 = 0.5f * (wW0 + wA0), wz1 = 0.5f * (wW1 + wA1);
            float fz0 = zf(uz[0], uz[1], uz[2], uz[3], wz0, k, nz, fnm, fnp);
            float fz1 = zf(uz[1], uz[2], uz[3], uz[4], wz1, k + 1, nz, fnm, fnp);
            float t = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy,
                           rdnw[k]);
            t = t - pgf_val(s.rdx, s.hrdx, c1, c2, rdnw[k], mux_c, dmu_u,
                            A.ph_0, A.ph_p, W.ph_0, W.ph_p, altA, altW,
                            A.pp_0, W.pp_0, alA, alW, pbA, pbW, B3,
                            phh(A), phh(W), A.dpn_0, A.dpn_p, W.dpn_0,
                            W.dpn_p);
            if (DIFF) {
                float fc = x[3];
                float lap = s.kh * (s.dxi2 * ((x[4] - 2.0f * fc) + x[2])
                                    + s.dyi2 * ((y[4] - 2.0f * fc) + y[2]));
                const float gup = (uz[3] - fc) * rdzf_h[k < nz - 1 ? k : nz - 2];
                const float gdn = (fc - uz[1]) * rdzf_h[k > 0 ? k - 1 : 0];
                float fup = (k < nz - 1) ? gup : 0.0f;
                float fdn = (k > 0) ? gdn : 0.0f;
                lap = lap + (s.kv * (fup - fdn)) * rdzc_h[k];
                t = t + (c1 * msf_u + c2) * lap;
            }
            gst(d.ru_t, lu + j * pu, cC, t);
            gst(d.u_pp, lu + j * pu, cC,
                (c1 * mtf_u + c2) * gld(d.u0, lu + j * pu, cC)
                    - (c1 * msf_u + c2) * x[3]);
            uz[0] = uz[1]; uz[1] = uz[2]; uz[2] = uz[3]; uz[3] = uz[4];
            const float un = gld(d.u, (k + 3 < nz ? k + 3 : nz - 1) * stu + j * pu, cC);
            uz[4] = (k + 3 < nz) ? un : 0.0f;
        }
        // ---- v face ----
        {
            float x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) y[o] = gld(d.v, lv + yo[o] * nx, cC);
#pragma unroll
            for (int o = 0; o < 7; ++o)
                x[o] = (o == 3) ? y[3] : gld(d.v, lv + j * nx, cx[o]);
            float ru_jS = (c1 * mux_jS + c2) * gld(d.u, lu + jm * pu, cC);
            float ru_jA = (c1 * mux_jA + c2) * gld(d.u, lu + j * pu, cC);
            float ru_pS = (c1 * mux_pS + c2) * gld(d.u, lu + jm * pu, ci1);
            float ru_pA = (c1 * mux_pA + c2) * gld(d.u, lu + j * pu, ci1);
            float u0c = 0.5f * (ru_jS + ru_jA), u1c = 0.5f * (ru_pS + ru_pA);
            float fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0c);
            float fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1c);
            const float v_b = gld(d.v, lv + jb * nx, cC);
            const float v_d = gld(d.v, lv + (j + 1) * nx, cC);
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
            t = t - pgf_val(s.rdy, s.hrdy, c1, c2, rdnw[k], muy_c, dmu_v,
                            A.ph_0, A.ph_p, S.ph_0, S.ph_p, altA, altS,
                            A.pp_0, S.pp_0, alA, alS, pbA, pbS, B3,
                            phh(A), phh(S), A.dpn_0, A.dpn_p, S.dpn_0,
                            S.dpn_p);
            if (DIFF) {
                float fc = y[3];
                float lap = s.kh * (s.dxi2 * ((x[4] - 2.0f * fc) + x[2])
                                    + s.dyi2 * ((y[4] - 2.0f * fc) + y[2]));
                const float gup = (vz[3] - fc) * rdzf_h[k < nz - 1 ? k : nz - 2];
                const float gdn = (fc - vz[1]) * rdzf_h[k > 0 ? k - 1 : 0];
                float fup = (k < nz - 1) ? gup : 0.0f;
                float fdn = (k > 0) ? gdn : 0.0f;
                lap = lap + (s.kv * (fup - fdn)) * rdzc_h[k];
                t = t + (c1 * msf_v + c2) * lap;
            }
            gst(d.rv_t, lv + j * nx, cC, t);
            gst(d.v_pp, lv + j * nx, cC,
                (c1 * mtf_v + c2) * gld(d.v0, lv + j * nx, cC)
                    - (c1 * msf_v + c2) * y[3]);
            vz[0] = vz[1]; vz[1] = vz[2]; vz[2] = vz[3]; vz[3] = vz[4];
            const float vn = gld(d.v, (k + 3 < nz ? k + 3 : nz - 1) * stv + j * nx, cC);
            vz[4] = (k + 3 < nz) ? vn : 0.0f;
        }
        wA0 = wA1; wW0 = wW1; wS0 = wS1;
        win_shift(A, k, rA, cC);
        win_shift(W, k, rA, cW);
        win_shift(S, k, rS, cC);
    }
}

// Periodic duplicate faces of the stage tendencies: the face-nx u tendency
// equals face 0's (same stencil inputs), the face-ny v tendency equals
// row 0's; u''/v'' init uses the duplicate's own u/u0 (v/v0).  One thread
// per (k, j) duplicate u face and per (k, i) duplicate v face.
__global__ __launch_bounds__(256) void k_tend_faces_edge(Dom d)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const int pu = nx + 1, stu = ny * pu, stv = (ny + 1) * nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const float* R mup0 = d.mup0;
    if (t < nz * ny) {
        const int k = t / ny, j = t - k * ny;
        const int cA = j * nx, cW = j * nx + nx - 1;
        const float mA = mub[cA] + mup[cA], mW = mub[cW] + mup[cW];
        const float mtA = mub[cA] + mup0[cA];
        const float mtf = 0.5f * (mtA + (mub[cW] + mup0[cW]));
        const float msf = 0.5f * (mA + mW);
        const float c1 = d.c1h[k], c2 = d.c2h[k];
        const int f0 = k * stu + j * pu, fd = f0 + nx;
        d.ru_t[fd] = d.ru_t[f0];
        d.u_pp[fd] = (c1 * mtf + c2) * d.u0[fd] - (c1 * msf + c2) * d.u[fd];
    } else if (t < nz * ny + nz * nx) {
        const int r = t - nz * ny;
        const int k = r / nx, i = r - k * nx;
        const int cA = i, cS = (ny - 1) * nx + i;
        const float mA = mub[cA] + mup[cA], mS = mub[cS] + mup[cS];
        const float mtA = mub[cA] + mup0[cA];
        const float mtf = 0.5f * (mtA + (mub[cS] + mup0[cS]));
        const float msf = 0.5f * (mA + mS);
        const float c1 = d.c1h[k], c2 = d.c2h[k];
        const int f0 = k * stv + i, fd = f0 + ny * nx;
        d.rv_t[fd] = d.rv_t[f0];
        d.v_pp[fd] = (c1 * mtf + c2) * d.v0[fd] - (c1 * msf + c2) * d.v[fd];
    }
}

// ---------------------------------------------------------------------------
// Slow tendencies in the mass column (theta, w, phi) + column init
// ---------------------------------------------------------------------------

// Slow tendencies in the mass column (j, i): rw_t (w advection, buoyancy
// and vertical PGF, diffusion; reference advection.cu:803-1051,
// dycore.cu:213-308, diffusion.cu:22-51), rph_t (WRF rhs_ph,
// dycore.cu:541-750), rth_t (theta advection + diffusion), and the
// w''/phi''/theta''/p''/alpha'' small-step initialisation
// (dycore.cu:1012-1085).  Row-segment waves; level kinds peel the surface
// (k = 0) and the lid (k = nz) so the interior level body is one block.
template <bool B3, bool DIFF, bool MOIST>
__global__ __launch_bounds__(256) void k_tend_cols(Dom d, Sc s)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int pu = nx + 1, stu = ny * pu, stv = (ny + 1) * nx;
    const RowSeg rs = row_seg(nx, ny);
    if (!rs.live) return;
    const int i = rs.i, j = rs.j;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    kfloat* c1f = kview(d.c1f);
    kfloat* c2f = kview(d.c2f);
    kfloat* fnm = kview(d.fnm);
    kfloat* fnp = kview(d.fnp);
    kfloat* rdn = kview(d.rdn);
    kfloat* rdnw = kview(d.rdnw);
    kfloat* thb1 = kview(d.thb);
    kfloat* phb1 = kview(d.phb);
    kfloat* pb1 = kview(d.pb);
    kfloat* rdzf_h = kview(d.rdzf_h);
    kfloat* rdzc_h = kview(d.rdzc_h);
    kfloat* rdzf_w = kview(d.rdzf_w);
    kfloat* rdzc_w = kview(d.rdzc_w);
    int yo[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) yo[o] = pmod(j + o - 3, ny) * nx;
    const int rA = j * nx, rS = yo[2], rN = yo[4];
    unsigned cx[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) cx[o] = 4u * pmod(i + o - 3, nx);
    unsigned ci1 = 4u * (i + 1);
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(cx[o]);
        colb_opaque(ci1);
    };
    unsigned& cC = cx[3];
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    auto mu2 = [&](int row, unsigned cb) {
        return gld(mub, row, cb) + gld(mup, row, cb);
    };
    const float mA = mu2(rA, cC), mW = mu2(rA, cx[2]);
    const float mE = mu2(rA, cx[4]), mS = mu2(rS, cC), mN = mu2(rN, cC);
    const float mux_i = 0.5f * (mA + mW), mux_ip = 0.5f * (mE + mA);
    const float muy_j = 0.5f * (mA + mS), muy_jp = 0.5f * (mN + mA);
    const float mupc = gld(mup, rA, cC);
    const float mt = gld(mub, rA, cC) + gld(d.mup0, rA, cC);
    const float mupp = gld(d.mup0, rA, cC) - mupc;
    gst(d.mu_pp, rA, cC, mupp);
    const float qx = s.qrdx, qy = s.qrdy;

    auto thb_at = [&](int k, int row, unsigned cb) {
        return B3 ? gld(d.thb, k * st + row, cb) : thb1[k];
    };
    auto phb_at = [&](int k, int row, unsigned cb) {
        return B3 ? gld(d.phb, k * st + row, cb) : phb1[k];
    };
    auto th_at = [&](int k, int row, unsigned cb) {
        return thb_at(k, row, cb) + gld(d.thp, k * st + row, cb);
    };
    auto ph_at = [&](int k, int row, unsigned cb) {
        return phb_at(k, row, cb) + gld(d.php, k * st + row, cb);
    };

    // windows
    float thz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};   // theta k-2..k+2
    float wz[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};    // w k-2..k+2
    for (int kk = 0; kk < 3; ++kk) {
        if (kk < nz) thz[2 + kk] = th_at(kk, rA, cC);
        wz[2 + kk] = gld(d.w, kk * st + rA, cC);
    }
    float ww_m = 0.0f, ww_0 = gld(d.ww, rA, cC), ww_p = gld(d.ww, st + rA, cC);
    float ph_m = 0.0f, ph_0 = ph_at(0, rA, cC), ph_p = ph_at(1, rA, cC);
    float pp_m = 0.0f;
    float ru_i_m = 0.0f, ru_ip_m = 0.0f, rv_j_m = 0.0f, rv_jp_m = 0.0f;
    float u_i_m = 0.0f, u_i_mm = 0.0f, u_ip_m = 0.0f, u_ip_mm = 0.0f;
    float v_j_m = 0.0f, v_j_mm = 0.0f, v_jp_m = 0.0f, v_jp_mm = 0.0f;
    float thp_m = 0.0f, thp_0 = gld(d.thp, rA, cC);
    float thp_p = gld(d.thp, st + rA, cC);
    float phpp_0 = gld(d.php0, rA, cC) - gld(d.php, rA, cC);
    // moist: total water (qv + qc) + qr at the level below (buoyancy)
    float qt_m = 0.0f, qt_mm = 0.0f;
    const float mubc = MOIST ? gld(mub, rA, cC) : 0.0f;

    // kinds: 0 = surface level (w level 0 + mass level 0), 1 = interior,
    // 2 = lid (w level nz only)
    auto level = [&](int k, auto kind) {
        constexpr int K = decltype(kind)::value;
        constexpr bool mass = K != 2;
        constexpr bool interior = K == 1;
        constexpr bool top = K == 2;
        opaque_all();
        const int lm = k * st;
        const int km = k < nz ? k : nz - 1;          // clamped mass level
        float pp_0 = 0.0f;
        float p_k = 0.0f;
        float qt_0 = 0.0f;
        if (mass) {
            p_k = gld(d.p, lm + rA, cC);
            pp_0 = p_k - (B3 ? gld(d.pb, lm + rA, cC) : pb1[k]);
            if (MOIST)
                qt_0 = (gld(d.qv, lm + rA, cC) + gld(d.qc, lm + rA, cC))
                       + gld(d.qr, lm + rA, cC);
        }
        float ru_i = 0.0f, ru_ip = 0.0f, rv_j = 0.0f, rv_jp = 0.0f;
        float u_i = 0.0f, u_ip = 0.0f, v_j = 0.0f, v_jp = 0.0f;
        if (mass) {
            const int lu = k * stu, lv = k * stv;
            u_i = gld(d.u, lu + j * pu, cC);
            const float u_i1 = gld(d.u, lu + j * pu, ci1);
            u_ip = gld(d.u, lu + j * pu, cx[4]);
            v_j = gld(d.v, lv + rA, cC);
            const float v_j1 = gld(d.v, lv + (j + 1) * nx, cC);
            v_jp = gld(d.v, lv + rN, cC);
            ru_i = (c1h[k] * mux_i + c2h[k]) * u_i;
            ru_ip = (c1h[k] * mux_ip + c2h[k]) * u_i1;
            rv_j = (c1h[k] * muy_j + c2h[k]) * v_j;
            rv_jp = (c1h[k] * muy_jp + c2h[k]) * v_j1;
        }

        // --- rw_t: advection + buoyancy + diffusion ---
        float t = 0.0f;
        float wx2 = 0.0f, wx4 = 0.0f, wy2 = 0.0f, wy4 = 0.0f;
        if (interior) {
            float x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) {
                x[o] = (o == 3) ? wz[2] : gld(d.w, lm + rA, cx[o]);
                y[o] = (o == 3) ? wz[2] : gld(d.w, lm + yo[o], cC);
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
                const int m = k - 1 + q;
                const float vel = 0.5f * (wwv[q] + wwv[q + 1]);
                const float lin = 0.5f * vel * (wz[q + 2] + wz[q + 1]);
                float cub = flux3(wz[q], wz[q + 1], wz[q + 2], wz[q + 3], vel);
                keep(cub);
                fz[q] = (m == 0 || m == nz - 1) ? lin : cub;
            }
            t = div3(fx1 - fx0, fy1 - fy0, fz[1] - fz[0], s.rdx, s.rdy, rdn[k]);
            float term;
            if (MOIST) {
                // moist pg_buoy_w (reference dycore.cu:279-289)
                const float cqw = 0.5f * (qt_0 + qt_m);
                const float cq1 = 1.0f / (1.0f + cqw);
                term = (cq1 * rdn[k]) * (pp_0 - pp_m) - c1f[k] * mupc;
                term = term - (cqw * cq1) * (c1f[k] * mubc + c2f[k]);
            } else {
                term = rdn[k] * (pp_0 - pp_m) - c1f[k] * mupc;
            }
            t = t + G_F * term;
        } else if (top) {
            if (MOIST) {
                // lid: the top two layers (reference dycore.cu:258-273);
                // qt_m holds layer nz-1, qt_mm layer nz-2
                const float cqw = 0.5f * (qt_m + qt_mm);
                const float cq1 = 1.0f / (1.0f + cqw);
                float pressure = ((cq1 * 2.0f) * rdnw[nz - 1]) * (0.0f - pp_m);
                float term = pressure - c1f[nz] * mupc;
                term = term - (cqw * cq1) * (c1f[nz] * mubc + c2f[nz]);
                t = t + G_F * term;
            } else {
                float pressure = ((1.0f * 2.0f) * rdnw[nz - 1]) * (0.0f - pp_m);
                float term = pressure - c1f[nz] * mupc;
                t = t + G_F * term;
            }
        }
        if (DIFF) {
            float lap = 0.0f;
            if (interior) {
                float fc = wz[2];
                lap = s.kh * (s.dxi2 * ((wx4 - 2.0f * fc) + wx2)
                              + s.dyi2 * ((wy4 - 2.0f * fc) + wy2));
                float fup = (wz[3] - fc) * rdzf_w[k];
                float fdn = (fc - wz[1]) * rdzf_w[k - 1];
                lap = lap + (s.kv * (fup - fdn)) * rdzc_w[k];
            }
            t = t + (c1f[k] * mA + c2f[k]) * lap;
        }
        gst(d.rw_t, lm + rA, cC, t);

        // --- rph_t: WRF rhs_ph ---
        float r = 0.0f;
        if (K != 0) {
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
            float fx = fcx_i * (p0 - ph_at(k, rA, cx[2]));
            float fxp = fcx_ip * (ph_at(k, rA, cx[4]) - p0);
            r = r - qxk * (fxp + fx);
            float fy = fcy_j * (p0 - ph_at(k, rS, cC));
            float fyp = fcy_jp * (ph_at(k, rN, cC) - p0);
            r = r - qyk * (fyp + fy);
        }
        gst(d.rph_t, lm + rA, cC, r);

        // --- w'' and phi'' initialisation at w level k ---
        gst(d.w_pp, lm + rA, cC,
            (c1f[k] * mt + c2f[k]) * gld(d.w0, lm + rA, cC)
                - (c1f[k] * mA + c2f[k]) * wz[2]);
        gst(d.ph_pp, lm + rA, cC, phpp_0);

        // --- mass level k: rth_t and the theta''/p''/alpha'' init ---
        float phpp_p = 0.0f;
        if (mass) {
            float qxv[7], qyv[7], tpx[7], tpy[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) {
                if (o == 3) continue;
                tpx[o] = gld(d.thp, lm + rA, cx[o]);
                tpy[o] = gld(d.thp, lm + yo[o], cC);
                qxv[o] = thb_at(k, rA, cx[o]) + tpx[o];
                qyv[o] = thb_at(k, yo[o], cC) + tpy[o];
            }
            qxv[3] = thz[2];
            qyv[3] = thz[2];
            float fx0 = flux5(qxv[0], qxv[1], qxv[2], qxv[3], qxv[4], qxv[5], ru_i);
            float fx1 = flux5(qxv[1], qxv[2], qxv[3], qxv[4], qxv[5], qxv[6], ru_ip);
            float fy0 = flux5(qyv[0], qyv[1], qyv[2], qyv[3], qyv[4], qyv[5], rv_j);
            float fy1 = flux5(qyv[1], qyv[2], qyv[3], qyv[4], qyv[5], qyv[6], rv_jp);
            float fz0 = zf(thz[0], thz[1], thz[2], thz[3], ww_0, k, nz, fnm, fnp);
            float fz1 = zf(thz[1], thz[2], thz[3], thz[4], ww_p, k + 1, nz, fnm, fnp);
            float t2 = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy, rdnw[k]);
            if (DIFF) {
                float fc = thp_0;
                float lap = s.kh * (s.dxi2 * ((tpx[4] - 2.0f * fc) + tpx[2])
                                    + s.dyi2 * ((tpy[4] - 2.0f * fc) + tpy[2]));
                const float gup = (thp_p - fc) * rdzf_h[k < nz - 1 ? k : nz - 2];
                const float gdn = (fc - thp_m) * rdzf_h[k > 0 ? k - 1 : 0];
                float fup = (k < nz - 1) ? gup : 0.0f;
                float fdn = (k > 0) ? gdn : 0.0f;
                lap = lap + (s.kv * (fup - fdn)) * rdzc_h[k];
                t2 = t2 + (c1h[k] * mA + c2h[k]) * lap;
            }
            gst(d.rth_t, lm + rA, cC, t2);

            phpp_p = gld(d.php0, lm + st + rA, cC) - gld(d.php, lm + st + rA, cC);
            float cht = c1h[k] * mt + c2h[k];
            float chs = c1h[k] * mA + c2h[k];
            float thbk = thb_at(k, rA, cC);
            float tht = thbk + gld(d.thp0, lm + rA, cC);
            float ths = thbk + thp_0;
            float thpp = cht * tht - chs * ths;
            gst(d.th_pp, lm + rA, cC, thpp);
            float altk = gld(d.alt, lm + rA, cC);
            float c2a = (GAMMA_F * p_k) / altk;
            float al = -((altk * (c1h[k] * mupp)) + rdnw[k] * (phpp_p - phpp_0)) / cht;
            if (B3) gst(d.al_pp, lm + rA, cC, al);
            float p5 = (altk * (thpp - (c1h[k] * mupp) * ths)) / (cht * ths);
            gst(d.p_pp, lm + rA, cC, c2a * (p5 - al));
        }
        (void)km;

        // shift windows (refills are clamped loads and selects)
        phpp_0 = phpp_p;
        const float th_n = th_at(k + 3 < nz ? k + 3 : nz - 1, rA, cC);
        thz[0] = thz[1]; thz[1] = thz[2]; thz[2] = thz[3]; thz[3] = thz[4];
        thz[4] = (k + 3 < nz) ? th_n : 0.0f;
        const float w_n = gld(d.w, (k + 3 <= nz ? k + 3 : nz) * st + rA, cC);
        wz[0] = wz[1]; wz[1] = wz[2]; wz[2] = wz[3]; wz[3] = wz[4];
        wz[4] = (k + 3 <= nz) ? w_n : 0.0f;
        const float ww_n = gld(d.ww, (k + 2 <= nz ? k + 2 : nz) * st + rA, cC);
        ww_m = ww_0; ww_0 = ww_p;
        ww_p = (k + 2 <= nz) ? ww_n : 0.0f;
        const float ph_n = ph_at(k + 2 <= nz ? k + 2 : nz, rA, cC);
        ph_m = ph_0; ph_0 = ph_p;
        ph_p = (k + 2 <= nz) ? ph_n : 0.0f;
        pp_m = pp_0;
        if (MOIST) {
            qt_mm = qt_m;
            qt_m = qt_0;
        }
        ru_i_m = ru_i; ru_ip_m = ru_ip; rv_j_m = rv_j; rv_jp_m = rv_jp;
        u_i_mm = u_i_m; u_i_m = u_i; u_ip_mm = u_ip_m; u_ip_m = u_ip;
        v_j_mm = v_j_m; v_j_m = v_j; v_jp_mm = v_jp_m; v_jp_m = v_jp;
        const float thp_n = gld(d.thp, (k + 2 < nz ? k + 2 : nz - 1) * st + rA, cC);
        thp_m = thp_0; thp_0 = thp_p;
        thp_p = (k + 2 < nz) ? thp_n : 0.0f;
    };
    level(0, Kind<0>{});
    for (int k = 1; k < nz; ++k) level(k, Kind<1>{});
    level(nz, Kind<2>{});
}

// ---------------------------------------------------------------------------
// Acoustic substep, part 1: u''/v'' forward step on faces
// ---------------------------------------------------------------------------

// Damped-pressure window of one column for the acoustic PGF.
// Loads of one u''/v'' level, issued one level ahead.
struct UvIn {
    float pA = 0.0f, oA = 0.0f, pW = 0.0f, oW = 0.0f, pS = 0.0f, oS = 0.0f;
    float phA = 0.0f, phW = 0.0f, phS = 0.0f;   // ph'' at w level k+1
    float hpA = 0.0f, hpW = 0.0f, hpS = 0.0f;   // php at w level k+1
    float gbA = 0.0f, gbW = 0.0f, gbS = 0.0f;   // phb at w level k+1 (terrain)
    float altA = 0.0f, altW = 0.0f, altS = 0.0f;
    float aA = 0.0f, aW = 0.0f, aS = 0.0f;      // al'' (terrain)
    float bA = 0.0f, bW = 0.0f, bS = 0.0f;      // pb (terrain)
    float u = 0.0f, ru = 0.0f, v = 0.0f, rv = 0.0f;
};

// u'' at face (j, iu) and/or v'' at face (jv, i) through all levels
// (reference acoustic.cu:244-314).  A is the mass column east of the u face
// and north of the v face, W its west and S its south neighbour.  The
// damped pressure p'' + smdiv (p'' - p''_old) rides in 4-level windows for
// the surface and lid extrapolations of dpn.
template <bool B3, bool USEP, bool DU, bool DV>
__device__ __forceinline__ void uv_march(const Dom& d, const Sc& s, int rA,
                                         int rS, int rU, unsigned bA,
                                         unsigned bW, unsigned bS, unsigned bu,
                                         unsigned bv)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const float smdiv = s.smdiv, dtau = s.dtau;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    kfloat* fnm = kview(d.fnm);
    kfloat* fnp = kview(d.fnp);
    kfloat* rdnw = kview(d.rdnw);
    auto opaque_all = [&]() {
        colb_opaque(bA); colb_opaque(bW); colb_opaque(bS);
        colb_opaque(bu); colb_opaque(bv);
    };
    auto pd = [&](float a, float o) {
        float b = USEP ? o : a;
        return a + smdiv * (a - b);
    };
    const float mA = gld(d.mub2d, rA, bA) + gld(d.mup, rA, bA);
    const float muf_u = DU ? 0.5f * ((gld(d.mub2d, rA, bW) + gld(d.mup, rA, bW)) + mA)
                           : 0.0f;
    const float muf_v = DV ? 0.5f * ((gld(d.mub2d, rS, bS) + gld(d.mup, rS, bS)) + mA)
                           : 0.0f;
    const float mppA = gld(d.mu_pp, rA, bA);
    const float dmu_u = DU ? gld(d.mu_pp, rA, bW) + mppA : 0.0f;
    const float dmu_v = DV ? gld(d.mu_pp, rS, bS) + mppA : 0.0f;

    // windows S(k-2), S(k-1), S(k), S(k+1) of the damped pressure
    opaque_all();
    float pa[4] = {0.0f, 0.0f, 0.0f, 0.0f}, pw[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float ps[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    pa[2] = pd(gld(d.p_pp, 0 + rA, bA), gld(d.p_pp_old, 0 + rA, bA));
    pa[3] = pd(gld(d.p_pp, st + rA, bA), gld(d.p_pp_old, st + rA, bA));
    const float pa2 = pd(gld(d.p_pp, 2 * st + rA, bA), gld(d.p_pp_old, 2 * st + rA, bA));
    float dnu_0 = 0.0f, dnv_0 = 0.0f;
    if (DU) {
        pw[2] = pd(gld(d.p_pp, 0 + rA, bW), gld(d.p_pp_old, 0 + rA, bW));
        pw[3] = pd(gld(d.p_pp, st + rA, bW), gld(d.p_pp_old, st + rA, bW));
        const float pw2 = pd(gld(d.p_pp, 2 * st + rA, bW), gld(d.p_pp_old, 2 * st + rA, bW));
        float S0 = pa[2] + pw[2], S1 = pa[3] + pw[3], S2 = pa2 + pw2;
        dnu_0 = 0.5f * ((d.cf1 * S0 + d.cf2 * S1) + d.cf3 * S2);
    }
    if (DV) {
        ps[2] = pd(gld(d.p_pp, 0 + rS, bS), gld(d.p_pp_old, 0 + rS, bS));
        ps[3] = pd(gld(d.p_pp, st + rS, bS), gld(d.p_pp_old, st + rS, bS));
        const float ps2 = pd(gld(d.p_pp, 2 * st + rS, bS), gld(d.p_pp_old, 2 * st + rS, bS));
        float S0 = pa[2] + ps[2], S1 = pa[3] + ps[3], S2 = pa2 + ps2;
        dnv_0 = 0.5f * ((d.cf1 * S0 + d.cf2 * S1) + d.cf3 * S2);
    }
    float phA0 = gld(d.ph_pp, 0 + rA, bA), hpA0 = gld(d.php, 0 + rA, bA);
    float phW0 = DU ? gld(d.ph_pp, 0 + rA, bW) : 0.0f, hpW0 = DU ? gld(d.php, 0 + rA, bW) : 0.0f;
    float phS0 = DV ? gld(d.ph_pp, 0 + rS, bS) : 0.0f, hpS0 = DV ? gld(d.php, 0 + rS, bS) : 0.0f;
    float gbA0 = B3 ? gld(d.phb, 0 + rA, bA) : 0.0f;
    float gbW0 = (B3 && DU) ? gld(d.phb, 0 + rA, bW) : 0.0f;
    float gbS0 = (B3 && DV) ? gld(d.phb, 0 + rS, bS) : 0.0f;

    auto load = [&](UvIn& L, int k) {
        const unsigned lp = (k + 2 < nz ? k + 2 : nz - 1) * st;
        const unsigned lw = (k + 1) * st, lm = k * st;
        L.pA = gld(d.p_pp, lp + rA, bA);
        if (USEP) L.oA = gld(d.p_pp_old, lp + rA, bA);
        L.phA = gld(d.ph_pp, lw + rA, bA);
        L.hpA = gld(d.php, lw + rA, bA);
        L.altA = gld(d.alt, lm + rA, bA);
        if (B3) {
            L.gbA = gld(d.phb, lw + rA, bA);
            L.aA = gld(d.al_pp, lm + rA, bA);
            L.bA = gld(d.pb, lm + rA, bA);
        }
        if (DU) {
            L.pW = gld(d.p_pp, lp + rA, bW);
            if (USEP) L.oW = gld(d.p_pp_old, lp + rA, bW);
            L.phW = gld(d.ph_pp, lw + rA, bW);
            L.hpW = gld(d.php, lw + rA, bW);
            L.altW = gld(d.alt, lm + rA, bW);
            if (B3) {
                L.gbW = gld(d.phb, lw + rA, bW);
                L.aW = gld(d.al_pp, lm + rA, bW);
                L.bW = gld(d.pb, lm + rA, bW);
            }
            L.u = gld(d.u_pp, k * stu + rU, bu);
            L.ru = gld(d.ru_t, k * stu + rU, bu);
        }
        if (DV) {
            L.pS = gld(d.p_pp, lp + rS, bS);
            if (USEP) L.oS = gld(d.p_pp_old, lp + rS, bS);
            L.phS = gld(d.ph_pp, lw + rS, bS);
            L.hpS = gld(d.php, lw + rS, bS);
            L.altS = gld(d.alt, lm + rS, bS);
            if (B3) {
                L.gbS = gld(d.phb, lw + rS, bS);
                L.aS = gld(d.al_pp, lm + rS, bS);
                L.bS = gld(d.pb, lm + rS, bS);
            }
            L.v = gld(d.v_pp, k * stv + rA, bv);
            L.rv = gld(d.rv_t, k * stv + rA, bv);
        }
    };

    UvIn nxt;
    load(nxt, 0);
    auto level = [&](int k, auto last_kind) {
        constexpr bool LAST = decltype(last_kind)::value;
        opaque_all();
        const UvIn L = nxt;
        if (!LAST) load(nxt, k + 1);
        const float c1 = c1h[k], c2 = c2h[k];
        const float phA1 = L.phA, hpA1 = L.hpA, gbA1 = L.gbA;
        const float altA = L.altA;
        auto dn_top = [&](float Sk, float Sm1, float Sm2) {
            return s.top_lid ? 0.5f * ((d.cf1 * Sk + d.cf2 * Sm1) + d.cf3 * Sm2)
                             : 0.0f;
        };
        if (DU) {
            float Sk = pa[2] + pw[2];
            float dn_up;
            if (LAST) {
                dn_up = dn_top(Sk, pa[1] + pw[1], pa[0] + pw[0]);
            } else {
                float Sp = pa[3] + pw[3];
                dn_up = 0.5f * (fnm[k + 1] * Sp + fnp[k + 1] * Sk);
            }
            float dph = (phA1 - L.phW) + (phA0 - phW0);
            float dpe = pa[2] - pw[2];
            float brk = dph + (altA + L.altW) * dpe;
            if (B3) brk = brk + (L.aA + L.aW) * (L.bA - L.bW);
            float dpxy = ((0.5f * s.rdx) * (c1 * muf_u + c2)) * brk;
            float phsA = hpA0 + hpA1, phsB = hpW0 + L.hpW;
            float dphp = 0.5f * (phsA - phsB);
            if (B3) dphp = dphp + 0.5f * ((gbA0 + gbA1) - (gbW0 + L.gbW));
            dpxy = dpxy + (s.rdx * dphp) * (rdnw[k] * (dn_up - dnu_0)
                                            - (0.5f * c1) * dmu_u);
            gst(d.u_pp, k * stu + rU, bu, L.u + dtau * (L.ru - dpxy));
            dnu_0 = dn_up;
            phW0 = L.phW; hpW0 = L.hpW; gbW0 = L.gbW;
        }
        if (DV) {
            float Sk = pa[2] + ps[2];
            float dn_up;
            if (LAST) {
                dn_up = dn_top(Sk, pa[1] + ps[1], pa[0] + ps[0]);
            } else {
                float Sp = pa[3] + ps[3];
                dn_up = 0.5f * (fnm[k + 1] * Sp + fnp[k + 1] * Sk);
            }
            float dph = (phA1 - L.phS) + (phA0 - phS0);
            float dpe = pa[2] - ps[2];
            float brk = dph + (altA + L.altS) * dpe;
            if (B3) brk = brk + (L.aA + L.aS) * (L.bA - L.bS);
            float dpxy = ((0.5f * s.rdy) * (c1 * muf_v + c2)) * brk;
            float phsA = hpA0 + hpA1, phsB = hpS0 + L.hpS;
            float dphp = 0.5f * (phsA - phsB);
            if (B3) dphp = dphp + 0.5f * ((gbA0 + gbA1) - (gbS0 + L.gbS));
            dpxy = dpxy + (s.rdy * dphp) * (rdnw[k] * (dn_up - dnv_0)
                                            - (0.5f * c1) * dmu_v);
            gst(d.v_pp, k * stv + rA, bv, L.v + dtau * (L.rv - dpxy));
            dnv_0 = dn_up;
            phS0 = L.phS; hpS0 = L.hpS; gbS0 = L.gbS;
        }
        phA0 = phA1; hpA0 = hpA1; gbA0 = gbA1;
        // advance the damped-pressure windows to k+1 (level k+2 enters)
        const bool in = k + 2 < nz;
        pa[0] = pa[1]; pa[1] = pa[2]; pa[2] = pa[3];
        pa[3] = in ? pd(L.pA, L.oA) : 0.0f;
        if (DU) {
            pw[0] = pw[1]; pw[1] = pw[2]; pw[2] = pw[3];
            pw[3] = in ? pd(L.pW, L.oW) : 0.0f;
        }
        if (DV) {
            ps[0] = ps[1]; ps[1] = ps[2]; ps[2] = ps[3];
            ps[3] = in ? pd(L.pS, L.oS) : 0.0f;
        }
    };
    for (int k = 0; k < nz - 1; ++k) level(k, std::false_type{});
    level(nz - 1, std::true_type{});
}

// One launch: the first ne workgroups advance the periodic duplicate faces
// (u'' at i = nx, one per row; v'' at j = ny, one per column; each with its
// own u''/ru_t and the PGF of face 0), dispatched first so they overlap the
// main work; the rest are row-segment waves in which the thread of mass
// column (j, i) advances u'' on face (j, i) and v'' on face (j, i).
template <bool B3, bool USEP>
__global__ __launch_bounds__(256) void k_uv(Dom d, Sc s, int ne)
{
    const int ny = d.ny, nx = d.nx, pu = nx + 1;
    if ((int)blockIdx.x < ne) {
        const int t = blockIdx.x * 256 + threadIdx.y * 64 + threadIdx.x;
        if (t < ny) {
            const int j = t;
            uv_march<B3, USEP, true, false>(
                d, s, 0, 0, 0, 4u * (j * nx), 4u * (j * nx + nx - 1), 0u,
                4u * (j * pu + nx), 0u);
        } else if (t < ny + nx) {
            const int i = t - ny;
            uv_march<B3, USEP, false, true>(
                d, s, 0, 0, 0, 4u * i, 0u, 4u * ((ny - 1) * nx + i), 0u,
                4u * (ny * nx + i));
        }
        return;
    }
    const RowSeg rs = row_seg_from(nx, ny, ne);
    if (!rs.live) return;
    const int i = rs.i, j = rs.j;
    const int jm = (j == 0) ? ny - 1 : j - 1;
    const unsigned bc = 4u * i, bw = 4u * ((i == 0) ? nx - 1 : i - 1);
    uv_march<B3, USEP, true, true>(d, s, j * nx, jm * nx, j * pu, bc, bw, bc,
                                   bc, bc);
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
// Inputs of upward level m (raw loads; all arithmetic on them happens in
// the iteration that consumes them, so the loads of level m+1 stay in
// flight while level m computes).
struct ColUp {
    float ui = 0.0f, uip = 0.0f;   // u'' at faces i, i+1
    float vj = 0.0f, vjp = 0.0f;   // v'' at faces j, j+1
    float te = 0.0f, tw = 0.0f, tn = 0.0f, ts = 0.0f;  // thp, 4 neighbours
    float be = 0.0f, bw = 0.0f, bn = 0.0f, bs = 0.0f;  // thb there (terrain)
    float thu = 0.0f, bhu = 0.0f;  // thp, thb at (m+1, c)
    float thpp = 0.0f, rth = 0.0f; // theta'' and its tendency
    float phu = 0.0f, ghu = 0.0f;  // php, phb at w level m+1
    float alt = 0.0f, p = 0.0f;    // stage reference alt, p
    float phpp = 0.0f, wpp = 0.0f, rph = 0.0f;  // at w level m
    float rw = 0.0f;               // w tendency at w level m-1
};

// Inputs of downward w level k (and mass level k when k < nz).
struct ColDn {
    float wpp = 0.0f, gam = 0.0f, w = 0.0f, php = 0.0f, phb = 0.0f;
    float phpp = 0.0f, thp = 0.0f, thb = 0.0f, alt = 0.0f, p = 0.0f;
    float thpp = 0.0f;
};


// B3: terrain (3-D) base state; DAMP: damp_opt = 3 w damper; AL: store
// alpha''; WW: store Omega'' (standalone substep API).  nz >= 3.
template <bool B3, bool DAMP, bool AL, bool WW>
__global__ __launch_bounds__(256) void k_column(Dom d, Sc s, float* R pout)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    const float eps = s.epssm, dtau = s.dtau;
    const float one_p = 1.0f + eps, one_m = 1.0f - eps;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    kfloat* c1f = kview(d.c1f);
    kfloat* c2f = kview(d.c2f);
    kfloat* dnw = kview(d.dnw);
    kfloat* rdnw = kview(d.rdnw);
    kfloat* rdn = kview(d.rdn);
    kfloat* fnm = kview(d.fnm);
    kfloat* fnp = kview(d.fnp);
    // per-thread column byte offsets
    unsigned bc = 4u * c;
    unsigned bE = 4u * (j * nx + ip), bW = 4u * (j * nx + im);
    unsigned bN = 4u * (jp * nx + i), bS = 4u * (jm * nx + i);
    unsigned bu = 4u * (j * (nx + 1) + i), bu1 = bu + 4u;
    unsigned bvn = 4u * (c + nx);
    auto opaque_all = [&]() {
        colb_opaque(bc); colb_opaque(bE); colb_opaque(bW);
        colb_opaque(bN); colb_opaque(bS); colb_opaque(bu);
        colb_opaque(bu1); colb_opaque(bvn);
    };

    const float mut = d.mub2d[c] + d.mup[c];
    const float mupp_old = d.mu_pp[c];
    float dmdt_pp = 0.0f;
#pragma unroll 5
    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const float a = gld(d.u_pp, k * stu, bu1), b = gld(d.u_pp, k * stu, bu);
        const float e = gld(d.v_pp, k * stv, bvn), f = gld(d.v_pp, k * stv, bc);
        dmdt_pp = dmdt_pp + dnw[k] * (s.rdx * (a - b) + s.rdy * (e - f));
    }
    const float rmu = d.rmu_t[c];
    const float mupp = mupp_old + dtau * ((dmdt_pp + d.dmdt_rf[c]) + rmu);
    d.mu_pp[c] = mupp;
    const float muts = mut + mupp;
    const float muave = 0.5f * (one_p * mupp + one_m * mupp_old);
    float cof = 0.5f * dtau * G_F * one_p;
    cof = cof * cof;
    auto cfm_at = [&](int kk) { return c1f[kk] * mut + c2f[kk]; };
    kfloat* thb1 = kview(d.thb);     // 1-D base profiles (flat base)
    kfloat* phb1 = kview(d.phb);
    auto thbv = [&](float loaded, int kk) { return B3 ? loaded : thb1[kk]; };
    auto phbv = [&](float loaded, int kk) { return B3 ? loaded : phb1[kk]; };

    // kinematic lower boundary on w'' from the updated u''/v''
    float w0;
    {
        opaque_all();
        auto lw = [&](const float* a, unsigned colb, int sl) {
            return (d.cf1 * gld(a, 0, colb) + d.cf2 * gld(a, sl, colb))
                   + d.cf3 * gld(a, 2 * sl, colb);
        };
        float ue = lw(d.u_pp, bu1, stu), uw = lw(d.u_pp, bu, stu);
        float vn = lw(d.v_pp, bvn, stv), vs = lw(d.v_pp, bc, stv);
        const float* R ht = d.ht;
        float h = ht[c];
        w0 = (0.5f * s.rdy) * ((ht[jp * nx + i] - h) * vn + (h - ht[jm * nx + i]) * vs)
             + (0.5f * s.rdx) * ((ht[j * nx + ip] - h) * ue + (h - ht[j * nx + im]) * uw);
        gst(d.w_pp, 0, bc, w0);
    }

    // q: level of the loads.  Mass and face loads clamp to nz-1 (their
    // values are unused past the top); w-level loads use q <= nz.
    auto load_up = [&](ColUp& L, int q) {
        const int qm = q < nz ? q : nz - 1;
        const int qu = (q + 1 < nz) ? q + 1 : nz - 1;
        const int qw = q + 1 <= nz ? q + 1 : nz;
        const unsigned lm = qm * st;
        L.ui = gld(d.u_pp, qm * stu, bu);
        L.uip = gld(d.u_pp, qm * stu, bu1);
        L.vj = gld(d.v_pp, qm * stv, bc);
        L.vjp = gld(d.v_pp, qm * stv, bvn);
        L.te = gld(d.thp, lm, bE);
        L.tw = gld(d.thp, lm, bW);
        L.tn = gld(d.thp, lm, bN);
        L.ts = gld(d.thp, lm, bS);
        L.thu = gld(d.thp, qu * st, bc);
        if (B3) {
            L.be = gld(d.thb, lm, bE);
            L.bw = gld(d.thb, lm, bW);
            L.bn = gld(d.thb, lm, bN);
            L.bs = gld(d.thb, lm, bS);
            L.bhu = gld(d.thb, qu * st, bc);
            L.ghu = gld(d.phb, qw * st, bc);
        }
        L.thpp = gld(d.th_pp, lm, bc);
        L.rth = gld(d.rth_t, lm, bc);
        L.phu = gld(d.php, qw * st, bc);
        L.alt = gld(d.alt, lm, bc);
        L.p = gld(d.p, lm, bc);
        L.phpp = gld(d.ph_pp, q * st, bc);
        L.wpp = gld(d.w_pp, q * st, bc);
        L.rph = gld(d.rph_t, q * st, bc);
        L.rw = gld(d.rw_t, (q >= 1 ? q - 1 : 0) * st, bc);
    };

    // upward sweep
    if (WW) gst(d.ww_pp, 0, bc, 0.0f);
    float ww_k = 0.0f, wwk = 0.0f, wdtn_k = 0.0f;
    float th_m = (B3 ? gld(d.thb, 0, bc) : thb1[0]) + gld(d.thp, 0, bc);
    float ph_m = (B3 ? gld(d.phb, 0, bc) : phb1[0]) + gld(d.php, 0, bc);
    float wd_prev = 0.0f;
    float rhs_a = 0.0f, rhs_b = 0.0f;
    float pho_a = 0.0f, pho_b = gld(d.ph_pp, 0, bc);
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

    auto force_row = [&](int r, float rwr, float rhs_dn, float rhs_c,
                         float rhs_up, float ph_dn, float ph_c, float ph_up,
                         float wold, float t2dn, float t2c, float c2dn,
                         float c2c, float altdn, float altc, float chmdn,
                         float chmc) {
        float dph_up = one_p * (rhs_up - rhs_c) + one_m * (ph_up - ph_c);
        float dph_dn = one_p * (rhs_c - rhs_dn) + one_m * (ph_c - ph_dn);
        float t1 = dtau * rwr
                   + (c_force * rdn[r])
                         * ((c2c * rdnw[r]) / chmc * dph_up
                            - (c2dn * rdnw[r - 1]) / chmdn * dph_dn);
        float t3 = c_buoy * (rdn[r] * ((c2c * altc) * t2c - (c2dn * altdn) * t2dn)
                             - c1f[r] * muave);
        return wold + (t1 + t3);
    };
    // LU factors of row r (1 <= r <= nz-1) from the windowed c2a/chm
    auto lu_row = [&](int r, float c2dn, float c2c, float chmdn, float chmc,
                      float& alpha_r, bool r_ge2) {
        float a_r = 0.0f;
        if (r_ge2)
            a_r = -cof * rdn[r] * rdnw[r - 1] * c2dn / (chmdn * cfm_at(r - 1));
        float b = 1.0f + cof * rdn[r] * (rdnw[r] * c2c / (chmc * cfm_at(r))
                                         + rdnw[r - 1] * c2dn / (chmdn * cfm_at(r)));
        float cc = -cof * rdn[r] * rdnw[r] * c2c / (chmc * cfm_at(r + 1));
        alpha_r = 1.0f / (b - a_r * gam_prev);
        gam_prev = cc * alpha_r;
        gst(d.gam, r * st, bc, gam_prev);
        return a_r;
    };

    ColUp nxt;
    load_up(nxt, 0);
    // One upward level; the kind fixes at compile time which parts exist:
    // rhs of phi'' from level 1, an eliminated row from level 2, and no
    // Omega/theta flux above the top level.
    auto up_level = [&](int m, auto kind) {
        constexpr int K = decltype(kind)::value;
        constexpr bool has_up = K != LV_LAST;       // m + 1 < nz
        constexpr bool has_rhs = K != LV_FIRST;     // m >= 1
        constexpr bool has_row = K >= LV_THIRD;     // m >= 2
        opaque_all();
        const ColUp L = nxt;
        load_up(nxt, m + 1);
        const unsigned ms = m * st;
        float dvp = s.rdx * (L.uip - L.ui) + s.rdy * (L.vjp - L.vj);
        float ww_kp1;
        if (has_up) {
            wwk = wwk - dnw[m] * (c1h[m] * (dmdt_pp + rmu) + dvp);
            ww_kp1 = wwk;
        } else {
            ww_kp1 = 0.0f;
        }
        if (WW) gst(d.ww_pp, ms + st, bc, ww_kp1);
        float th_up = has_up ? thbv(L.bhu, m + 1) + L.thu : 0.0f;
        float wdtn_kp1 = has_up
            ? ww_kp1 * (fnm[m + 1] * th_up + fnp[m + 1] * th_m) : 0.0f;
        float t_c = th_m;
        float hx = (0.5f * s.rdx) * (L.uip * (thbv(L.be, m) + L.te + t_c)
                                     - L.ui * (t_c + (thbv(L.bw, m) + L.tw)));
        float hy = (0.5f * s.rdy) * (L.vjp * (thbv(L.bn, m) + L.tn + t_c)
                                     - L.vj * (t_c + (thbv(L.bs, m) + L.ts)));
        float tho = L.thpp;
        float thn = tho + dtau * (L.rth - ((hx + hy) + rdnw[m] * (wdtn_kp1 - wdtn_k)));
        gst(d.th_pp, ms, bc, thn);
        float chs = c1h[m] * muts + c2h[m];
        float t2 = (0.5f * (one_p * thn + one_m * tho)) / (chs * t_c);
        float ph_up = phbv(L.ghu, m + 1) + L.phu;
        float wd = ((0.5f * (ww_kp1 + ww_k)) * rdnw[m]) * (ph_up - ph_m);
        float altm = L.alt;
        float c2a = GAMMA_F * L.p / altm;
        float chm = c1h[m] * mut + c2h[m];
        float rhs_c = 0.0f, pho_c = pho_b, wold_c = 0.0f;
        if (has_rhs) {
            pho_c = L.phpp;
            wold_c = L.wpp;
            float r = dtau * (L.rph + ghalf_eps * wold_c);
            r = r - dtau * (fnm[m] * wd + fnp[m] * wd_prev);
            rhs_c = pho_c + r / cfm_at(m);
            gst(d.ph_pp, ms, bc, rhs_c);
        }
        if (has_row) {
            const int r = m - 1;
            float alpha_r;
            const bool r_ge2 = K == LV_MID ? true
                               : K == LV_THIRD ? false : r >= 2;
            float a_r = lu_row(r, c2a_a, c2a_b, chm_a, chm_b, alpha_r, r_ge2);
            float wr = force_row(r, L.rw, rhs_a, rhs_b, rhs_c, pho_a, pho_b,
                                 pho_c, wold_b, t2_a, t2_b, c2a_a, c2a_b,
                                 alt_a, alt_b, chm_a, chm_b);
            float wf = (wr - a_r * wf_prev) * alpha_r;
            gst(d.w_pp, r * st, bc, wf);
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
    };
    up_level(0, Kind<LV_FIRST>{});
    up_level(1, Kind<LV_SECOND>{});
    if (nz > 3) up_level(2, Kind<LV_THIRD>{});
    for (int m = 3; m < nz - 1; ++m) up_level(m, Kind<LV_MID>{});
    up_level(nz - 1, Kind<LV_LAST>{});
    {
        // nxt holds the w-level nz loads (phi'', w'', rph_t) and rw_t at nz-1
        const unsigned f = nz * st;
        const float pho_c = nxt.phpp;
        const float wold_c = nxt.wpp;
        float r = dtau * (nxt.rph + ghalf_eps * wold_c);
        float rhs_c = pho_c + r / cfm_at(nz);
        if (s.top_lid) rhs_c = 0.0f;
        gst(d.ph_pp, f, bc, rhs_c);
        const int rr = nz - 1;
        float alpha_r;
        float a_r = lu_row(rr, c2a_a, c2a_b, chm_a, chm_b, alpha_r, rr >= 2);
        float wr = force_row(rr, nxt.rw, rhs_a, rhs_b, rhs_c, pho_a, pho_b,
                             pho_c, wold_b, t2_a, t2_b, c2a_a, c2a_b, alt_a,
                             alt_b, chm_a, chm_b);
        float wfr = (wr - a_r * wf_prev) * alpha_r;
        gst(d.w_pp, rr * st, bc, wfr);
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
            wtop = wold_c + ((dtau * d.rw_t[f + c] - bterm) - cterm);
        }
        float b_top = 1.0f + 2.0f * cof * rdnw[nz - 1] * rdnw[nz - 1] * c2a_b
                             / (chm_b * cfm_at(nz));
        float alpha_top = 1.0f / (b_top - a_top * gam_prev);
        gst(d.w_pp, f, bc, (wtop - a_top * wfr) * alpha_top);
    }

    // downward sweep: back substitution, damper, phi'' update, EOS
    float htop = 0.0f, hbot = 0.0f;
    if (DAMP) {
        htop = ((B3 ? d.phb[nz * st + c] : phb1[nz]) + d.php[nz * st + c]) / G_F;
        hbot = htop - s.zdamp;
    }
    const float c_ph = ((0.5f * dtau) * G_F) * one_p;
    auto load_dn = [&](ColDn& L, int k) {
        const unsigned ks = k * st;
        const unsigned km = (k < nz ? k : nz - 1) * st;
        L.wpp = gld(d.w_pp, ks, bc);
        L.gam = gld(d.gam, ks, bc);
        L.phpp = gld(d.ph_pp, ks, bc);
        if (DAMP) {
            L.w = gld(d.w, ks, bc);
            L.php = gld(d.php, ks, bc);
            if (B3) L.phb = gld(d.phb, ks, bc);
        }
        L.thp = gld(d.thp, km, bc);
        if (B3) L.thb = gld(d.thb, km, bc);
        L.alt = gld(d.alt, km, bc);
        L.p = gld(d.p, km, bc);
        L.thpp = gld(d.th_pp, km, bc);
    };
    float wb_next = 0.0f, phn_up = 0.0f;
    ColDn dn;
    load_dn(dn, nz);
    // kinds: LV_FIRST = top w level (no mass level, no gamma), LV_MID,
    // LV_LAST = surface (phi'' only)
    auto dn_level = [&](int k, auto kind) {
        constexpr int K = decltype(kind)::value;
        colb_opaque(bc);
        const ColDn L = dn;
        load_dn(dn, k >= 1 ? k - 1 : 0);
        const unsigned f = k * st;
        float phn;
        if (K != LV_LAST) {
            float wb = L.wpp;
            if (K != LV_FIRST) wb = wb - L.gam * wb_next;
            wb_next = wb;
            float wfin = wb;
            if (DAMP) {
                float hk = (phbv(L.phb, k) + L.php) / G_F;
                if (hk >= hbot) {
                    float arg = (1.5707963f * (hk - hbot)) / s.zdamp;
                    float sn = (float)sin((double)arg);
                    float dw = (s.dampmag * sn) * sn;
                    wfin = (wb - (dw * cfm_at(k)) * L.w) / (1.0f + dw);
                }
            }
            gst(d.w_pp, f, bc, wfin);
            float cfs = c1f[k] * muts + c2f[k];
            phn = L.phpp + (c_ph * wfin) / cfs;
            gst(d.ph_pp, f, bc, phn);
        } else {
            phn = L.phpp;
        }
        if (K != LV_FIRST) {
            float chm2 = c1h[k] * muts + c2h[k];
            float th_ref = thbv(L.thb, k) + L.thp;
            float altk = L.alt;
            float c2a = GAMMA_F * L.p / altk;
            float al = -((altk * (c1h[k] * mupp)) + rdnw[k] * (phn_up - phn)) / chm2;
            if (AL) gst(d.al_pp, f, bc, al);
            gst(pout, f, bc, c2a * ((altk * (L.thpp - (c1h[k] * mupp) * th_ref))
                                    / (chm2 * th_ref) - al));
        }
        phn_up = phn;
    };
    dn_level(nz, Kind<LV_FIRST>{});
    for (int k = nz - 1; k >= 1; --k) dn_level(k, Kind<LV_MID>{});
    dn_level(0, Kind<LV_LAST>{});
}

// ---------------------------------------------------------------------------
// Small-step finish (reference dycore.cu:1087-1198) + EOS of the new state
// ---------------------------------------------------------------------------

// Coupled-to-uncoupled fold of one face (reference dycore.cu:1087-1127):
// new = (C(mu_t*) x_t* + x'') / C(mu_t* + mu'').
template <int STRIDE_KIND>
__device__ __forceinline__ void finish_face(const Dom& d, float* R out,
                                            const float* R x,
                                            const float* R xpp, int f,
                                            float msf, float mnf)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const int sf = STRIDE_KIND == 0 ? ny * (nx + 1) : (ny + 1) * nx;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    unsigned bf = 4u * f;
    for (int k = 0; k < nz; ++k) {
        colb_opaque(bf);
        const float c1 = c1h[k], c2 = c2h[k];
        gst(out, k * sf, bf, ((c1 * msf + c2) * gld(x, k * sf, bf)
                              + gld(xpp, k * sf, bf)) / (c1 * mnf + c2));
    }
}

struct FinIn {
    float u = 0.0f, upp = 0.0f, v = 0.0f, vpp = 0.0f;
    float w = 0.0f, wpp = 0.0f, php = 0.0f, phpp = 0.0f;
    float thp = 0.0f, thpp = 0.0f;
    float thb = 0.0f, phb = 0.0f, alb = 0.0f;   // terrain base
};

// Small-step finish (reference dycore.cu:1087-1198) of the mass column
// (j, i), its u face (j, i) and v face (j, i), then the EOS of the new
// state (diagnostics.cu:37-89).
template <bool B3, bool EOS>
__global__ __launch_bounds__(256) void k_finish(Dom d, FinishDst o)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int im = (i == 0) ? nx - 1 : i - 1;
    const int jm = (j == 0) ? ny - 1 : j - 1;
    kfloat* c1h = kview(d.c1h);
    kfloat* c2h = kview(d.c2h);
    kfloat* c1f = kview(d.c1f);
    kfloat* c2f = kview(d.c2f);
    kfloat* rdnw = kview(d.rdnw);
    kfloat* thb1 = kview(d.thb);
    kfloat* phb1 = kview(d.phb);
    kfloat* alb1 = kview(d.alb);
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const float* R mu_pp = d.mu_pp;
    const float msA = mub[c] + mup[c], mnA = msA + mu_pp[c];
    const int cW = j * nx + im, cS = jm * nx + i;
    const float msW = mub[cW] + mup[cW], mnW = msW + mu_pp[cW];
    const float msS = mub[cS] + mup[cS], mnS = msS + mu_pp[cS];
    const float msf_u = 0.5f * (msA + msW), mnf_u = 0.5f * (mnA + mnW);
    const float msf_v = 0.5f * (msA + msS), mnf_v = 0.5f * (mnA + mnS);
    const float mup_new = mup[c] + mu_pp[c];
    const float mu_eos = mub[c] + mup_new;
    unsigned bc = 4u * c, bu = 4u * (j * (nx + 1) + i);

    auto load = [&](FinIn& L, int k) {
        const int km = k < nz ? k : nz - 1;
        L.u = gld(d.u, km * stu, bu);
        L.upp = gld(d.u_pp, km * stu, bu);
        L.v = gld(d.v, km * stv, bc);
        L.vpp = gld(d.v_pp, km * stv, bc);
        L.w = gld(d.w, k * st, bc);
        L.wpp = gld(d.w_pp, k * st, bc);
        L.php = gld(d.php, k * st, bc);
        L.phpp = gld(d.ph_pp, k * st, bc);
        L.thp = gld(d.thp, km * st, bc);
        L.thpp = gld(d.th_pp, km * st, bc);
        if (B3) {
            L.thb = gld(d.thb, km * st, bc);
            L.phb = gld(d.phb, k * st, bc);
            L.alb = gld(d.alb, km * st, bc);
        }
    };
    float php_lo = 0.0f, thp_lo = 0.0f, thb_lo = 0.0f, phb_lo = 0.0f;
    float alb_lo = 0.0f;
    FinIn nxt;
    load(nxt, 0);
    // kinds: 0 = surface (no EOS below), 1 = interior, 2 = top w level
    auto level = [&](int k, auto kind) {
        constexpr int K = decltype(kind)::value;
        colb_opaque(bc);
        colb_opaque(bu);
        const FinIn L = nxt;
        if (K != 2) load(nxt, k + 1);
        const unsigned f = k * st;
        if (K != 2) {
            const float c1 = c1h[k], c2 = c2h[k];
            gst(o.u, k * stu, bu, ((c1 * msf_u + c2) * L.u + L.upp)
                                      / (c1 * mnf_u + c2));
            gst(o.v, k * stv, bc, ((c1 * msf_v + c2) * L.v + L.vpp)
                                      / (c1 * mnf_v + c2));
        }
        gst(o.w, f, bc, ((c1f[k] * msA + c2f[k]) * L.w + L.wpp)
                            / (c1f[k] * mnA + c2f[k]));
        const float php_new = L.php + L.phpp;
        gst(o.php, f, bc, php_new);
        const float phb_k = B3 ? L.phb : phb1[k];
        if (EOS && K != 0) {
            const unsigned h = f - st;
            float th = thb_lo + thp_lo;
            float dph = (phb_k + php_new) - (phb_lo + php_lo);
            float alt = (-dph * rdnw[k - 1]) / (c1h[k - 1] * mu_eos + c2h[k - 1]);
            gst(d.alt, h, bc, alt);
            gst(d.al, h, bc, alt - alb_lo);
            gst(d.p, h, bc, P0_F * pow_gamma((RD_F * th) / (P0_F * alt)));
        }
        if (K != 2) {
            const float thbk = B3 ? L.thb : thb1[k];
            float num = (c1h[k] * msA + c2h[k]) * (thbk + L.thp) + L.thpp;
            float thn = num / (c1h[k] * mnA + c2h[k]) - thbk;
            gst(o.thp, f, bc, thn);
            thp_lo = thn;
            thb_lo = thbk;
            alb_lo = B3 ? L.alb : alb1[k];
        }
        php_lo = php_new;
        phb_lo = phb_k;
    };
    level(0, Kind<0>{});
    for (int k = 1; k < nz; ++k) level(k, Kind<1>{});
    level(nz, Kind<2>{});
    o.mup[c] = mup_new;
}

// Periodic duplicate faces of the finish: u at i = nx, v at j = ny.
__global__ __launch_bounds__(256) void k_finish_edge(Dom d, FinishDst o)
{
    const int ny = d.ny, nx = d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float* R mub = d.mub2d;
    const float* R mup = d.mup;
    const float* R mu_pp = d.mu_pp;
    if (t < ny) {
        const int j = t, cA = j * nx, cB = j * nx + nx - 1;
        const float msA = mub[cA] + mup[cA], mnA = msA + mu_pp[cA];
        const float msB = mub[cB] + mup[cB], mnB = msB + mu_pp[cB];
        finish_face<0>(d, o.u, d.u, d.u_pp, j * (nx + 1) + nx,
                       0.5f * (msA + msB), 0.5f * (mnA + mnB));
    } else if (t < ny + nx) {
        const int i = t - ny, cA = i, cB = (ny - 1) * nx + i;
        const float msA = mub[cA] + mup[cA], mnA = msA + mu_pp[cA];
        const float msB = mub[cB] + mup[cB], mnB = msB + mu_pp[cB];
        finish_face<1>(d, o.v, d.v, d.v_pp, ny * nx + i,
                       0.5f * (msA + msB), 0.5f * (mnA + mnB));
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline dim3 cols(const Dom& d)
{
    return dim3(xcd_grid((d.ny * d.nx + 255) / 256));
}
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
    const bool b3 = d.base3d != 0;
    if (do_eos && b3)
        hipLaunchKernelGGL((k_omega<true, true>), cols(d), dim3(256), 0, st, d, s);
    else if (do_eos)
        hipLaunchKernelGGL((k_omega<true, false>), cols(d), dim3(256), 0, st, d, s);
    else
        hipLaunchKernelGGL((k_omega<false, false>), cols(d), dim3(256), 0, st, d, s);
    return hipGetLastError();
}

hipError_t run_tend_faces(const Dom& d, const Sc& s, hipStream_t st)
{
    const dim3 g = rowseg_grid(d.nx, d.ny), b(64, 4);
    if (d.base3d && s.diff)
        hipLaunchKernelGGL((k_tend_faces<true, true>), g, b, 0, st, d, s);
    else if (d.base3d)
        hipLaunchKernelGGL((k_tend_faces<true, false>), g, b, 0, st, d, s);
    else if (s.diff)
        hipLaunchKernelGGL((k_tend_faces<false, true>), g, b, 0, st, d, s);
    else
        hipLaunchKernelGGL((k_tend_faces<false, false>), g, b, 0, st, d, s);
    const int n = d.nz * d.ny + d.nz * d.nx;
    hipLaunchKernelGGL(k_tend_faces_edge, dim3((n + 255) / 256), dim3(256), 0,
                       st, d);
    return hipGetLastError();
}

template <bool MOIST>
static void launch_tend_cols(const Dom& d, const Sc& s, hipStream_t st)
{
    const dim3 g = rowseg_grid(d.nx, d.ny), b(64, 4);
    if (d.base3d && s.diff)
        hipLaunchKernelGGL((k_tend_cols<true, true, MOIST>), g, b, 0, st, d, s);
    else if (d.base3d)
        hipLaunchKernelGGL((k_tend_cols<true, false, MOIST>), g, b, 0, st, d, s);
    else if (s.diff)
        hipLaunchKernelGGL((k_tend_cols<false, true, MOIST>), g, b, 0, st, d, s);
    else
        hipLaunchKernelGGL((k_tend_cols<false, false, MOIST>), g, b, 0, st, d, s);
}

hipError_t run_tend_cols(const Dom& d, const Sc& s, hipStream_t st)
{
    if (d.qv) launch_tend_cols<true>(d, s, st);
    else launch_tend_cols<false>(d, s, st);
    return hipGetLastError();
}

template <bool B3, bool USEP>
static void launch_uv(const Dom& d, const Sc& s, hipStream_t st)
{
    const int ne = (int)xcd_grid((unsigned)((d.ny + d.nx + 255) / 256));
    const dim3 g(rowseg_grid(d.nx, d.ny).x + ne);
    hipLaunchKernelGGL((k_uv<B3, USEP>), g, dim3(64, 4), 0, st, d, s, ne);
}

hipError_t run_uv(const Dom& d, const Sc& s, hipStream_t st)
{
    const bool b3 = d.base3d != 0, usep = s.use_pold != 0;
    if (b3 && usep) launch_uv<true, true>(d, s, st);
    else if (b3) launch_uv<true, false>(d, s, st);
    else if (usep) launch_uv<false, true>(d, s, st);
    else launch_uv<false, false>(d, s, st);
    return hipGetLastError();
}

template <bool B3, bool DAMP, bool AL, bool WW>
static void launch_column(const Dom& d, const Sc& s, float* pout,
                          hipStream_t st)
{
    hipLaunchKernelGGL((k_column<B3, DAMP, AL, WW>), cols(d), dim3(256), 0, st,
                       d, s, pout);
}

template <bool B3, bool DAMP>
static void launch_column_bd(const Dom& d, const Sc& s, float* pout, int al,
                             int ww, hipStream_t st)
{
    if (al && ww) launch_column<B3, DAMP, true, true>(d, s, pout, st);
    else if (al) launch_column<B3, DAMP, true, false>(d, s, pout, st);
    else if (ww) launch_column<B3, DAMP, false, true>(d, s, pout, st);
    else launch_column<B3, DAMP, false, false>(d, s, pout, st);
}

hipError_t run_column(const Dom& d, const Sc& s, float* pout, int write_al,
                      int write_ww, hipStream_t st)
{
    const bool b3 = d.base3d != 0, damp = s.damp != 0;
    if (b3 && damp) launch_column_bd<true, true>(d, s, pout, write_al, write_ww, st);
    else if (b3) launch_column_bd<true, false>(d, s, pout, write_al, write_ww, st);
    else if (damp) launch_column_bd<false, true>(d, s, pout, write_al, write_ww, st);
    else launch_column_bd<false, false>(d, s, pout, write_al, write_ww, st);
    return hipGetLastError();
}

hipError_t run_finish(const Dom& d, const FinishDst& o, hipStream_t st)
{
    // a moist state's EOS waits for the scalar update of the stage
    const bool eos = d.qv == nullptr;
    if (d.base3d && eos)
        hipLaunchKernelGGL((k_finish<true, true>), cols(d), dim3(256), 0, st, d, o);
    else if (d.base3d)
        hipLaunchKernelGGL((k_finish<true, false>), cols(d), dim3(256), 0, st, d, o);
    else if (eos)
        hipLaunchKernelGGL((k_finish<false, true>), cols(d), dim3(256), 0, st, d, o);
    else
        hipLaunchKernelGGL((k_finish<false, false>), cols(d), dim3(256), 0, st, d, o);
    hipLaunchKernelGGL(k_finish_edge, dim3((d.ny + d.nx + 255) / 256),
                       dim3(256), 0, st, d, o);
    return hipGetLastError();
}
