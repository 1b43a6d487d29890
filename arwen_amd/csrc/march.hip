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

namespace ARW_NS {

#define RS __restrict__

namespace {

// Compile-time level kinds: loops peel their special levels so the
// interior level body has no branches.
template <int V>
using Kind = std::integral_constant<int, V>;
enum { LV_FIRST = 0, LV_SECOND = 1, LV_THIRD = 2, LV_MID = 3, LV_LAST = 4 };

__device__ __forceinline__ real mu_at(const real* RS mub, const real* RS mup,
                                       int c)
{
    return mub[c] + mup[c];
}

// Vertical face flux of a half-level field at w-level kf from the window
// q(kf-2), q(kf-1), q(kf), q(kf+1) (reference advection.cu:189-201).
template <class CP>
__device__ __forceinline__ real zf(real qm2, real qm1, real q0, real qp1,
                                    real vel, int kf, int nz, CP fnm, CP fnp)
{
    // both forms are evaluated and one is selected, so a level loop built
    // on zf stays one basic block (fnm/fnp have nz entries)
    const int kc = kf < nz ? kf : nz - 1;
    const real lin = vel * (fnm[kc] * q0 + fnp[kc] * qm1);
    real cub = flux3(qm2, qm1, q0, qp1, vel);
    keep(cub);
    const real r = (kf == 1 || kf == nz - 1) ? lin : cub;
    return (kf == 0 || kf == nz) ? R(0.0) : r;
}

__device__ __forceinline__ real div3(real dfx, real dfy, real dfz,
                                      real rdx, real rdy, real rdz)
{
    real t = -dfx * rdx;
    t = t - dfy * rdy;
    return t - dfz * rdz;
}

// Large-step PGF on a face (reference dycore.cu:98-157) from values at
// level k of the two columns A (right/north) and B.
__device__ __forceinline__ real pgf_val(
    real rd, real hrd, real c1, real c2, real rdnwk, real muf, real dmu,
    real phA0, real phA1, real phB0, real phB1, real altA, real altB,
    real ppA, real ppB, real alA, real alB, real pbA, real pbB,
    int base3d, real phhA, real phhB, real dpnA0, real dpnA1, real dpnB0,
    real dpnB1)
{
    real layer = c1 * muf + c2;
    real bracket = (phA1 - phB1) + (phA0 - phB0);
    bracket = bracket + (altA + altB) * (ppA - ppB);
    if (base3d)
        bracket = bracket + (alA + alB) * (pbA - pbB);
    else
        bracket = bracket + (alA + alB) * R(0.0);
    real left = (hrd * layer) * bracket;
    real dphp = phhA - phhB;
    real hi = R(0.5) * (dpnA1 + dpnB1);
    real lo = R(0.5) * (dpnA0 + dpnB0);
    real vertical = rdnwk * (hi - lo) - c1 * dmu;
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
// calc_p_alpha (diagnostics.cu:37-89) when the stage state is new; in the
// precision form (PREC) its perturbation form, eos_pert, which also keeps
// p' in d.pp.
template <class SRC, bool EOS, bool B3, bool PREC, bool XO = true>
__global__ __launch_bounds__(256) void k_omega(typename SRC::Arg a, Sc s)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const int c = (SRC::many ? g : xcd_block()) * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* dnw = kview(d.dnw);
    kreal* rdnw = kview(d.rdnw);
    unsigned bc = RB * c, bu = RB * (j * (nx + 1) + i), bu1 = bu + RB;
    unsigned bvn = RB * (c + nx);
    const real ms = d.mub2d[c] + d.mup[c];
    if (EOS) {
        kreal* thb1 = kview(d.thb);
        kreal* phb1 = kview(d.phb);
        kreal* alb1 = kview(d.alb);
        kreal* pb1 = kview(d.pb);
        const real mupc = d.mup[c];
        real ph_lo = gld(d.php, 0, bc);
        real gb_lo = B3 ? gld(d.phb, 0, bc) : phb1[0];
#pragma unroll 2
        for (int k = 0; k < nz; ++k) {
            colb_opaque(bc);
            const unsigned h = k * st;
            const real ph_hi = gld(d.php, h + st, bc);
            const real gb_hi = B3 ? gld(d.phb, h + st, bc) : phb1[k + 1];
            const real albk = B3 ? gld(d.alb, h, bc) : alb1[k];
            if (PREC) {
                const real thbk = B3 ? gld(d.thb, h, bc) : thb1[k];
                const real pbk = B3 ? gld(d.pb, h, bc) : pb1[k];
                const real qvk = d.qv ? gld(d.qv, h, bc) : R(0.0);
                const EosPert e = eos_pert(gld(d.thp, h, bc), thbk, qvk,
                                           d.qv != nullptr, albk, pbk, mupc,
                                           c1h[k] * ms + c2h[k], c1h[k],
                                           rdnw[k], ph_hi - ph_lo);
                gst(d.alt, h, bc, e.alt);
                gst(d.al, h, bc, e.al);
                gst(d.p, h, bc, e.p);
                gst(d.pp, h, bc, e.pp);
            } else {
                real th = (B3 ? gld(d.thb, h, bc) : thb1[k]) + gld(d.thp, h, bc);
                if (d.qv) th = th * (R(1.0) + RVOVRD_F * gld(d.qv, h, bc));
                real dph = (gb_hi + ph_hi) - (gb_lo + ph_lo);
                real alt, al;
                eos_alpha<XO>(d, k, ms, dph, albk, c1h[k], c2h[k], rdnw[k], alt,
                          al);
                gst(d.alt, h, bc, alt);
                gst(d.al, h, bc, al);
                gst(d.p, h, bc, P0_F * pow_gamma((RD_F * th) / (P0_F * alt)));
            }
            ph_lo = ph_hi;
            gb_lo = gb_hi;
        }
    }
    const real mw = d.mub2d[j * nx + im] + d.mup[j * nx + im];
    const real me = d.mub2d[j * nx + ip] + d.mup[j * nx + ip];
    const real mS = d.mub2d[jm * nx + i] + d.mup[jm * nx + i];
    const real mN = d.mub2d[jp * nx + i] + d.mup[jp * nx + i];
    const real muxw = R(0.5) * (ms + mw), muxe = R(0.5) * (me + ms);
    const real muys = R(0.5) * (ms + mS), muyn = R(0.5) * (mN + ms);
    real dmdt = R(0.0), rf = R(0.0);
    Kahan ksum;
#pragma unroll 4
    for (int k = 0; k < nz; ++k) {
        colb_opaque(bc); colb_opaque(bu); colb_opaque(bu1); colb_opaque(bvn);
        const real c1 = c1h[k], c2 = c2h[k];
        real ruw = (c1 * muxw + c2) * gld(d.u, k * stu, bu);
        real rue = (c1 * muxe + c2) * gld(d.u, k * stu, bu1);
        real rvs = (c1 * muys + c2) * gld(d.v, k * stv, bc);
        real rvn = (c1 * muyn + c2) * gld(d.v, k * stv, bvn);
        real divv = dnw[k] * (s.rdx * (rue - ruw) + s.rdy * (rvn - rvs));
        if (PREC) {
            ksum.add(divv);
        } else {
            dmdt = (k == 0) ? divv : dmdt + divv;
            rf = rf + divv;
        }
        gst(d.ww, (k + 1) * st, bc, divv);
    }
    if (PREC) {
        // precision form: the compensated column sum serves both
        dmdt = ksum.s;
        rf = ksum.s;
    }
    gst(d.ww, 0, bc, R(0.0));
    real run = R(0.0);
    Kahan krun;
#pragma unroll 4
    for (int k = 0; k < nz - 1; ++k) {
        colb_opaque(bc);
        const unsigned h = (k + 1) * st;
        real inc = (c1h[k] * dnw[k]) * dmdt + gld(d.ww, h, bc);
        if (PREC) {
            krun.add(inc);
            run = krun.s;
        } else {
            run = run + inc;
        }
        gst(d.ww, h, bc, -run);
    }
    gst(d.ww, nz * st, bc, R(0.0));
    d.dmdt_rf[c] = rf;
    d.rmu_t[c] = R(0.0);
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

// Row segment of a member's logical workgroup L (batched launches).
__device__ __forceinline__ RowSeg row_seg_at(int nx, int ny, int L)
{
    const int nbx = (nx + 63) / 64;
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
    real pp_m = R(0.0), pp_0 = R(0.0), pp_p = R(0.0);
    real ph_0 = R(0.0), ph_p = R(0.0), gb_0 = R(0.0), gb_p = R(0.0);
    real dpn_0 = R(0.0), dpn_p = R(0.0);
};

// Slow tendencies on the u face (j, i) and v face (j, i): 5th/3rd-order
// advection (reference advection.cu:424-794), large-step PGF
// (dycore.cu:98-157), constant-K diffusion (diffusion.cu:22-51), and the
// u''/v'' small-step initialisation (dycore.cu:947-1010).  The periodic
// duplicate faces are finished by k_tend_faces_edge.  PREC (precision
// form): the pressure gradient takes p' from the perturbation-form EOS.
template <class SRC, bool B3, bool DIFF, bool PREC, bool HELD = false,
          bool XO = true>
__global__ __launch_bounds__(256) void k_tend_faces(typename SRC::Arg a, Sc s)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int pu = nx + 1;                       // u row pitch
    const int stu = ny * pu, stv = (ny + 1) * nx;
    const RowSeg rs = SRC::many ? row_seg_at(nx, ny, g) : row_seg(nx, ny);
    if (!rs.live) return;
    const int i = rs.i, j = rs.j;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    kreal* rdnw = kview(d.rdnw);
    kreal* rdzf_h = kview(d.rdzf_h);
    kreal* rdzc_h = kview(d.rdzc_h);
    // wave-uniform rows
    int yo[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) yo[o] = pmod(j + o - 3, ny);
    const int jm = yo[2], jp = yo[4];
    const int jb = (j == 0) ? ny : j;            // v_b row (duplicate at j=0)
    // per-lane column byte offsets
    unsigned cx[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) cx[o] = RB * pmod(i + o - 3, nx);
    unsigned ci1 = RB * (i + 1);                 // unwrapped east face
    unsigned cub = RB * ((i == 0) ? nx : i);     // u_b column
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(cx[o]);
        colb_opaque(ci1); colb_opaque(cub);
    };
    unsigned& cC = cx[3];                        // centre, west
    unsigned& cW = cx[2];
    // mass columns: A = (j, i), W = (j, i-1), S = (j-1, i)
    const int rA = j * nx, rS = jm * nx;
    const real* RS mub = d.mub2d;
    const real* RS mup = d.mup;
    const real* RS mup0 = d.mup0;
    auto mu2 = [&](int row, unsigned cb) {
        return gld(mub, row, cb) + gld(mup, row, cb);
    };
    const real mA = mu2(rA, cC), mW = mu2(rA, cW), mS = mu2(rS, cC);
    const real mE = mu2(rA, cx[4]), mN = mu2(jp * nx, cC);
    // u-face masses
    const real mux_a = R(0.5) * (mW + mu2(rA, cx[1]));
    const real mux_c = R(0.5) * (mA + mW);
    const real mux_d = R(0.5) * (mE + mA);
    const real muy_jW = R(0.5) * (mW + mu2(rS, cW));
    const real muy_jA = R(0.5) * (mA + mS);
    const real muy_pW = R(0.5) * (mu2(jp * nx, cW) + mW);
    const real muy_pA = R(0.5) * (mN + mA);
    // v-face masses
    const real muy_a = R(0.5) * (mS + mu2(yo[1] * nx, cC));
    const real muy_c = R(0.5) * (mA + mS);
    const real muy_d = R(0.5) * (mN + mA);
    const real mux_jS = R(0.5) * (mS + mu2(rS, cW));
    const real mux_jA = R(0.5) * (mA + mW);
    const real mux_pS = R(0.5) * (mu2(rS, cx[4]) + mS);
    const real mux_pA = R(0.5) * (mE + mA);
    // PGF and init masses
    const real dmu_u = R(0.5) * (gld(mup, rA, cC) + gld(mup, rA, cW));
    const real dmu_v = R(0.5) * (gld(mup, rA, cC) + gld(mup, rS, cC));
    const real mtA = gld(mub, rA, cC) + gld(mup0, rA, cC);
    const real mtf_u = R(0.5) * (mtA + (gld(mub, rA, cW) + gld(mup0, rA, cW)));
    const real mtf_v = R(0.5) * (mtA + (gld(mub, rS, cC) + gld(mup0, rS, cC)));
    const real msf_u = R(0.5) * (mA + mW), msf_v = R(0.5) * (mA + mS);

    // PGF windows of A, W, S
    auto pp_at = [&](int k, int row, unsigned cb) {
        const int l = k * st + row;
        if (PREC) return gld(d.pp, l, cb);      // the EOS's own p'
        return gld(d.p, l, cb) - (B3 ? gld(d.pb, l, cb) : kview(d.pb)[k]);
    };
    auto win_init = [&](PgfWin& X, int row, unsigned cb) {
        real p0 = pp_at(0, row, cb), p1 = pp_at(1, row, cb);
        real p2 = pp_at(2, row, cb);
        X.pp_0 = p0;
        X.pp_p = p1;
        X.ph_0 = gld(d.php, row, cb);
        X.gb_0 = B3 ? gld(d.phb, row, cb) : R(0.0);
        X.dpn_0 = (d.cf1 * p0 + d.cf2 * p1) + d.cf3 * p2;
    };
    auto win_load = [&](PgfWin& X, int k, int row, unsigned cb, bool top) {
        X.ph_p = gld(d.php, (k + 1) * st + row, cb);
        X.gb_p = B3 ? gld(d.phb, (k + 1) * st + row, cb) : R(0.0);
        const int kp = top ? nz - 1 : k + 1;
        const real lid = s.top_lid ? d.cfn * X.pp_0 + d.cfn1 * X.pp_m : R(0.0);
        const real mid = fnm[kp] * X.pp_p + fnp[kp] * X.pp_0;
        X.dpn_p = top ? lid : mid;
    };
    auto win_shift = [&](PgfWin& X, int k, int row, unsigned cb) {
        X.pp_m = X.pp_0;
        X.pp_0 = X.pp_p;
        const real np = pp_at(k + 2 < nz ? k + 2 : nz - 1, row, cb);
        X.pp_p = (k + 2 < nz) ? np : R(0.0);
        X.ph_0 = X.ph_p;
        X.gb_0 = X.gb_p;
        X.dpn_0 = X.dpn_p;
    };
    auto phh = [&](const PgfWin& X) {
        if (B3) return R(0.5) * ((X.gb_0 + X.gb_p) + (X.ph_0 + X.ph_p));
        return R(0.5) * (X.ph_0 + X.ph_p);
    };
    PgfWin A, W, S;
    win_init(A, rA, cC);
    win_init(W, rA, cW);
    win_init(S, rS, cC);

    real uz[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
    real vz[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
    for (int kk = 0; kk < 3 && kk < nz; ++kk) {
        uz[2 + kk] = gld(d.u, kk * stu + j * pu, cC);
        vz[2 + kk] = gld(d.v, kk * stv + j * nx, cC);
    }
    // the vertical flux through w level k+1 is the next level's flux
    // through its lower face: carried, not recomputed (level 0's lower
    // face flux is zero)
    real fzu = R(0.0), fzv = R(0.0);

    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const bool top = k + 1 == nz;
        const real c1 = c1h[k], c2 = c2h[k];
        const int lm = k * st;
        win_load(A, k, rA, cC, top);
        win_load(W, k, rA, cW, top);
        win_load(S, k, rS, cC, top);
        const real wA1 = gld(d.ww, lm + st + rA, cC);
        const real wW1 = gld(d.ww, lm + st + rA, cW);
        const real wS1 = gld(d.ww, lm + st + rS, cC);
        const real altA = gld(d.alt, lm + rA, cC), alA = gld(d.al, lm + rA, cC);
        const real altW = gld(d.alt, lm + rA, cW), alW = gld(d.al, lm + rA, cW);
        const real altS = gld(d.alt, lm + rS, cC), alS = gld(d.al, lm + rS, cC);
        const real pbA = B3 ? gld(d.pb, lm + rA, cC) : R(0.0);
        const real pbW = B3 ? gld(d.pb, lm + rA, cW) : R(0.0);
        const real pbS = B3 ? gld(d.pb, lm + rS, cC) : R(0.0);
        const int lu = k * stu, lv = k * stv;
        // ---- u face ----
        {
            real x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) x[o] = gld(d.u, lu + j * pu, cx[o]);
#pragma unroll
            for (int o = 0; o < 7; ++o)
                y[o] = (o == 3) ? x[3] : gld(d.u, lu + yo[o] * pu, cC);
            const real u_b = gld(d.u, lu + j * pu, cub);
            const real u_d = gld(d.u, lu + j * pu, ci1);
            real ru_a = (c1 * mux_a + c2) * x[2];
            real ru_b = (c1 * mux_c + c2) * u_b;
            real ru_c = (c1 * mux_c + c2) * x[3];
            real ru_d = (c1 * mux_d + c2) * u_d;
            real vl = R(0.5) * (ru_a + ru_b), vr = R(0.5) * (ru_c + ru_d);
            real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vl);
            real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], vr);
            real rv_jW = (c1 * muy_jW + c2) * gld(d.v, lv + j * nx, cW);
            real rv_jA = (c1 * muy_jA + c2) * gld(d.v, lv + j * nx, cC);
            real rv_pW = (c1 * muy_pW + c2) * gld(d.v, lv + (j + 1) * nx, cW);
            real rv_pA = (c1 * muy_pA + c2) * gld(d.v, lv + (j + 1) * nx, cC);
            real v0 = R(0.5) * (rv_jW + rv_jA), v1 = R(0.5) * (rv_pW + rv_pA);
            real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0);
            real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1);
            const real wz1 = R(0.5) * (wW1 + wA1);
            const real fz0 = fzu;
            const real fz1 = zf(uz[1], uz[2], uz[3], uz[4], wz1, k + 1, nz, fnm, fnp);
            fzu = fz1;
            real t = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy,
                           rdnw[k]);
            real pgu = pgf_val(s.rdx, s.hrdx, c1, c2, rdnw[k], mux_c, dmu_u,
                               A.ph_0, A.ph_p, W.ph_0, W.ph_p, altA, altW,
                               A.pp_0, W.pp_0, alA, alW, pbA, pbW, B3,
                               phh(A), phh(W), A.dpn_0, A.dpn_p, W.dpn_0,
                               W.dpn_p);
            if (XO && s.cq) pgu = gld(d.cqu, lu + j * pu, cC) * pgu;   // moist_cq
            t = t - pgu;
            if (DIFF) {
                real fc = x[3];
                real lap = s.kh * (s.dxi2 * ((x[4] - R(2.0) * fc) + x[2])
                                    + s.dyi2 * ((y[4] - R(2.0) * fc) + y[2]));
                const real gup = (uz[3] - fc) * rdzf_h[k < nz - 1 ? k : nz - 2];
                const real gdn = (fc - uz[1]) * rdzf_h[k > 0 ? k - 1 : 0];
                real fup = (k < nz - 1) ? gup : R(0.0);
                real fdn = (k > 0) ? gdn : R(0.0);
                lap = lap + (s.kv * (fup - fdn)) * rdzc_h[k];
                t = t + (c1 * msf_u + c2) * lap;
            }
            // the held mixing tendency comes last (reference dycore.py:
            // 2389-2390, add_fixed_dry_tendencies)
            if (HELD) t = t + gld(d.d6_u, lu + j * pu, cC);
            gst(d.ru_t, lu + j * pu, cC, t);
            gst(d.u_pp, lu + j * pu, cC,
                (c1 * mtf_u + c2) * gld(d.u0, lu + j * pu, cC)
                    - (c1 * msf_u + c2) * x[3]);
            uz[0] = uz[1]; uz[1] = uz[2]; uz[2] = uz[3]; uz[3] = uz[4];
            const real un = gld(d.u, (k + 3 < nz ? k + 3 : nz - 1) * stu + j * pu, cC);
            uz[4] = (k + 3 < nz) ? un : R(0.0);
        }
        // ---- v face ----
        {
            real x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) y[o] = gld(d.v, lv + yo[o] * nx, cC);
#pragma unroll
            for (int o = 0; o < 7; ++o)
                x[o] = (o == 3) ? y[3] : gld(d.v, lv + j * nx, cx[o]);
            real ru_jS = (c1 * mux_jS + c2) * gld(d.u, lu + jm * pu, cC);
            real ru_jA = (c1 * mux_jA + c2) * gld(d.u, lu + j * pu, cC);
            real ru_pS = (c1 * mux_pS + c2) * gld(d.u, lu + jm * pu, ci1);
            real ru_pA = (c1 * mux_pA + c2) * gld(d.u, lu + j * pu, ci1);
            real u0c = R(0.5) * (ru_jS + ru_jA), u1c = R(0.5) * (ru_pS + ru_pA);
            real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0c);
            real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1c);
            const real v_b = gld(d.v, lv + jb * nx, cC);
            const real v_d = gld(d.v, lv + (j + 1) * nx, cC);
            real rv_a = (c1 * muy_a + c2) * y[2];
            real rv_b = (c1 * muy_c + c2) * v_b;
            real rv_c = (c1 * muy_c + c2) * y[3];
            real rv_d = (c1 * muy_d + c2) * v_d;
            real vl = R(0.5) * (rv_a + rv_b), vr = R(0.5) * (rv_c + rv_d);
            real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], vl);
            real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], vr);
            const real wz1 = R(0.5) * (wS1 + wA1);
            const real fz0 = fzv;
            const real fz1 = zf(vz[1], vz[2], vz[3], vz[4], wz1, k + 1, nz, fnm, fnp);
            fzv = fz1;
            real t = div3(fx1 - fx0, fy1 - fy0, fz1 - fz0, s.rdx, s.rdy,
                           rdnw[k]);
            real pgv = pgf_val(s.rdy, s.hrdy, c1, c2, rdnw[k], muy_c, dmu_v,
                               A.ph_0, A.ph_p, S.ph_0, S.ph_p, altA, altS,
                               A.pp_0, S.pp_0, alA, alS, pbA, pbS, B3,
                               phh(A), phh(S), A.dpn_0, A.dpn_p, S.dpn_0,
                               S.dpn_p);
            if (XO && s.cq) pgv = gld(d.cqv, lv + j * nx, cC) * pgv;    // moist_cq
            t = t - pgv;
            if (DIFF) {
                real fc = y[3];
                real lap = s.kh * (s.dxi2 * ((x[4] - R(2.0) * fc) + x[2])
                                    + s.dyi2 * ((y[4] - R(2.0) * fc) + y[2]));
                const real gup = (vz[3] - fc) * rdzf_h[k < nz - 1 ? k : nz - 2];
                const real gdn = (fc - vz[1]) * rdzf_h[k > 0 ? k - 1 : 0];
                real fup = (k < nz - 1) ? gup : R(0.0);
                real fdn = (k > 0) ? gdn : R(0.0);
                lap = lap + (s.kv * (fup - fdn)) * rdzc_h[k];
                t = t + (c1 * msf_v + c2) * lap;
            }
            if (HELD) t = t + gld(d.d6_v, lv + j * nx, cC);
            gst(d.rv_t, lv + j * nx, cC, t);
            gst(d.v_pp, lv + j * nx, cC,
                (c1 * mtf_v + c2) * gld(d.v0, lv + j * nx, cC)
                    - (c1 * msf_v + c2) * y[3]);
            vz[0] = vz[1]; vz[1] = vz[2]; vz[2] = vz[3]; vz[3] = vz[4];
            const real vn = gld(d.v, (k + 3 < nz ? k + 3 : nz - 1) * stv + j * nx, cC);
            vz[4] = (k + 3 < nz) ? vn : R(0.0);
        }
        win_shift(A, k, rA, cC);
        win_shift(W, k, rA, cW);
        win_shift(S, k, rS, cC);
    }
}

// Periodic duplicate faces of the stage tendencies: the face-nx u tendency
// equals face 0's (same stencil inputs), the face-ny v tendency equals
// row 0's; u''/v'' init uses the duplicate's own u/u0 (v/v0).  One thread
// per (k, j) duplicate u face and per (k, i) duplicate v face.
template <class SRC>
__global__ __launch_bounds__(256) void k_tend_faces_edge(typename SRC::Arg a)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const int pu = nx + 1, stu = ny * pu, stv = (ny + 1) * nx;
    const int t = g * blockDim.x + threadIdx.x;
    const real* RS mub = d.mub2d;
    const real* RS mup = d.mup;
    const real* RS mup0 = d.mup0;
    if (t < nz * ny) {
        const int k = t / ny, j = t - k * ny;
        const int cA = j * nx, cW = j * nx + nx - 1;
        const real mA = mub[cA] + mup[cA], mW = mub[cW] + mup[cW];
        const real mtA = mub[cA] + mup0[cA];
        const real mtf = R(0.5) * (mtA + (mub[cW] + mup0[cW]));
        const real msf = R(0.5) * (mA + mW);
        const real c1 = d.c1h[k], c2 = d.c2h[k];
        const int f0 = k * stu + j * pu, fd = f0 + nx;
        d.ru_t[fd] = d.ru_t[f0];
        d.u_pp[fd] = (c1 * mtf + c2) * d.u0[fd] - (c1 * msf + c2) * d.u[fd];
    } else if (t < nz * ny + nz * nx) {
        const int r = t - nz * ny;
        const int k = r / nx, i = r - k * nx;
        const int cA = i, cS = (ny - 1) * nx + i;
        const real mA = mub[cA] + mup[cA], mS = mub[cS] + mup[cS];
        const real mtA = mub[cA] + mup0[cA];
        const real mtf = R(0.5) * (mtA + (mub[cS] + mup0[cS]));
        const real msf = R(0.5) * (mA + mS);
        const real c1 = d.c1h[k], c2 = d.c2h[k];
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
// PREC (precision form): buoyancy takes p' from the perturbation-form EOS;
// rhs_ph's layer thickness is dphb + dphp with dphb = -(dnw (c1h mub +
// c2h)) alb (the base state's hydrostatic relation) and, on a flat base,
// its horizontal differences are those of php; theta'' is formed as
// (c1h mu'') thb + (C(mu_t) thp_t - C(mu_s) thp_s), free of the base
// theta's cancellation.
template <class SRC, bool B3, bool DIFF, bool MOIST, bool PREC,
          bool HELD = false, bool XO = true>
__global__ __launch_bounds__(256) void k_tend_cols(typename SRC::Arg a, Sc s)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int pu = nx + 1, stu = ny * pu, stv = (ny + 1) * nx;
    const RowSeg rs = SRC::many ? row_seg_at(nx, ny, g) : row_seg(nx, ny);
    if (!rs.live) return;
    const int i = rs.i, j = rs.j;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* c1f = kview(d.c1f);
    kreal* c2f = kview(d.c2f);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    kreal* rdn = kview(d.rdn);
    kreal* rdnw = kview(d.rdnw);
    kreal* thb1 = kview(d.thb);
    kreal* phb1 = kview(d.phb);
    kreal* pb1 = kview(d.pb);
    kreal* alb1 = kview(d.alb);
    kreal* dnw = kview(d.dnw);
    kreal* rdzf_h = kview(d.rdzf_h);
    kreal* rdzc_h = kview(d.rdzc_h);
    kreal* rdzf_w = kview(d.rdzf_w);
    kreal* rdzc_w = kview(d.rdzc_w);
    int yo[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) yo[o] = pmod(j + o - 3, ny) * nx;
    const int rA = j * nx, rS = yo[2], rN = yo[4];
    unsigned cx[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) cx[o] = RB * pmod(i + o - 3, nx);
    unsigned ci1 = RB * (i + 1);
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(cx[o]);
        colb_opaque(ci1);
    };
    unsigned& cC = cx[3];
    const real* RS mub = d.mub2d;
    const real* RS mup = d.mup;
    auto mu2 = [&](int row, unsigned cb) {
        return gld(mub, row, cb) + gld(mup, row, cb);
    };
    const real mA = mu2(rA, cC), mW = mu2(rA, cx[2]);
    const real mE = mu2(rA, cx[4]), mS = mu2(rS, cC), mN = mu2(rN, cC);
    const real mux_i = R(0.5) * (mA + mW), mux_ip = R(0.5) * (mE + mA);
    const real muy_j = R(0.5) * (mA + mS), muy_jp = R(0.5) * (mN + mA);
    const real mupc = gld(mup, rA, cC);
    const real mt = gld(mub, rA, cC) + gld(d.mup0, rA, cC);
    const real mupp = gld(d.mup0, rA, cC) - mupc;
    gst(d.mu_pp, rA, cC, mupp);
    const real qx = s.qrdx, qy = s.qrdy;

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
    // the geopotential the vertical window carries: full, or php alone in
    // the precision form (whose rhs_ph takes the base part from alb)
    auto phw_at = [&](int k, int row, unsigned cb) {
        if (PREC) return gld(d.php, k * st + row, cb);
        return ph_at(k, row, cb);
    };
    // horizontal neighbours of rhs_ph: php alone on a flat base in the
    // precision form (phb is horizontally uniform there)
    auto phh_at = [&](int k, int row, unsigned cb) {
        if (PREC && !B3) return gld(d.php, k * st + row, cb);
        return ph_at(k, row, cb);
    };

    // windows
    real thz[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};   // theta k-2..k+2
    real wz[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};    // w k-2..k+2
    for (int kk = 0; kk < 3; ++kk) {
        if (kk < nz) thz[2 + kk] = th_at(kk, rA, cC);
        wz[2 + kk] = gld(d.w, kk * st + rA, cC);
    }
    real ww_m = R(0.0), ww_0 = gld(d.ww, rA, cC), ww_p = gld(d.ww, st + rA, cC);
    real ph_m = R(0.0), ph_0 = phw_at(0, rA, cC), ph_p = phw_at(1, rA, cC);
    real pp_m = R(0.0);
    real ru_i_m = R(0.0), ru_ip_m = R(0.0), rv_j_m = R(0.0), rv_jp_m = R(0.0);
    real u_i_m = R(0.0), u_i_mm = R(0.0), u_ip_m = R(0.0), u_ip_mm = R(0.0);
    real v_j_m = R(0.0), v_j_mm = R(0.0), v_jp_m = R(0.0), v_jp_mm = R(0.0);
    real thp_m = R(0.0), thp_0 = gld(d.thp, rA, cC);
    real thp_p = gld(d.thp, st + rA, cC);
    real phpp_0 = gld(d.php0, rA, cC) - gld(d.php, rA, cC);
    // moist: total water (qv + qc) + qr at the level below (buoyancy)
    real qt_m = R(0.0), qt_mm = R(0.0);
    const real mubc = (MOIST || PREC) ? gld(mub, rA, cC) : R(0.0);
    // precision form: base layer thickness at mass level k
    auto dphb = [&](int k) {
        const real albk = B3 ? gld(d.alb, k * st + rA, cC) : alb1[k];
        return -(dnw[k] * (c1h[k] * mubc + c2h[k])) * albk;
    };

    // vertical fluxes through the upper face of the level below, carried
    // up the column: theta's at w level k, w's at mass level k-1
    real fzt = R(0.0), fzw = R(0.0);

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
        real pp_0 = R(0.0);
        real p_k = R(0.0);
        real qt_0 = R(0.0);
        if (mass) {
            p_k = gld(d.p, lm + rA, cC);
            if (PREC)
                pp_0 = gld(d.pp, lm + rA, cC);
            else
                pp_0 = p_k - (B3 ? gld(d.pb, lm + rA, cC) : pb1[k]);
            if (MOIST) {
                qt_0 = (gld(d.qv, lm + rA, cC) + gld(d.qc, lm + rA, cC))
                       + gld(d.qr, lm + rA, cC);
                // a mixed-phase scheme's condensate masses, in order
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (XO && e < d.n_load) qt_0 = qt_0 + gld(d.qx[e], lm + rA, cC);
            }
        }
        real ru_i = R(0.0), ru_ip = R(0.0), rv_j = R(0.0), rv_jp = R(0.0);
        real u_i = R(0.0), u_ip = R(0.0), v_j = R(0.0), v_jp = R(0.0);
        if (mass) {
            const int lu = k * stu, lv = k * stv;
            u_i = gld(d.u, lu + j * pu, cC);
            const real u_i1 = gld(d.u, lu + j * pu, ci1);
            u_ip = gld(d.u, lu + j * pu, cx[4]);
            v_j = gld(d.v, lv + rA, cC);
            const real v_j1 = gld(d.v, lv + (j + 1) * nx, cC);
            v_jp = gld(d.v, lv + rN, cC);
            ru_i = (c1h[k] * mux_i + c2h[k]) * u_i;
            ru_ip = (c1h[k] * mux_ip + c2h[k]) * u_i1;
            rv_j = (c1h[k] * muy_j + c2h[k]) * v_j;
            rv_jp = (c1h[k] * muy_jp + c2h[k]) * v_j1;
        }

        // --- rw_t: advection + buoyancy + diffusion ---
        real t = R(0.0);
        real wx2 = R(0.0), wx4 = R(0.0), wy2 = R(0.0), wy4 = R(0.0);
        if (interior) {
            real x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) {
                x[o] = (o == 3) ? wz[2] : gld(d.w, lm + rA, cx[o]);
                y[o] = (o == 3) ? wz[2] : gld(d.w, lm + yo[o], cC);
            }
            wx2 = x[2]; wx4 = x[4]; wy2 = y[2]; wy4 = y[4];
            real u0v = fnm[k] * ru_i + fnp[k] * ru_i_m;
            real u1v = fnm[k] * ru_ip + fnp[k] * ru_ip_m;
            real v0v = fnm[k] * rv_j + fnp[k] * rv_j_m;
            real v1v = fnm[k] * rv_jp + fnp[k] * rv_jp_m;
            real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], u0v);
            real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], u1v);
            real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], v0v);
            real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], v1v);
            // flux through mass level k (between w levels k and k+1)
            const real vel = R(0.5) * (ww_0 + ww_p);
            const real lin = R(0.5) * vel * (wz[3] + wz[2]);
            real cub = flux3(wz[1], wz[2], wz[3], wz[4], vel);
            keep(cub);
            const real fz1 = (k == nz - 1) ? lin : cub;
            t = div3(fx1 - fx0, fy1 - fy0, fz1 - fzw, s.rdx, s.rdy, rdn[k]);
            fzw = fz1;
            real term;
            if (MOIST) {
                // moist pg_buoy_w (reference dycore.cu:279-289)
                const real cqw = R(0.5) * (qt_0 + qt_m);
                const real cq1 = R(1.0) / (R(1.0) + cqw);
                term = (cq1 * rdn[k]) * (pp_0 - pp_m) - c1f[k] * mupc;
                term = term - (cqw * cq1) * (c1f[k] * mubc + c2f[k]);
            } else {
                term = rdn[k] * (pp_0 - pp_m) - c1f[k] * mupc;
            }
            t = t + G_F * term;
        } else if (!top) {
            // surface: the flux through mass level 0, which w level 1
            // takes as its lower one (the linear form)
            const real vel = R(0.5) * (ww_0 + ww_p);
            fzw = R(0.5) * vel * (wz[3] + wz[2]);
        } else if (top) {
            if (MOIST) {
                // lid: the top two layers (reference dycore.cu:258-273);
                // qt_m holds layer nz-1, qt_mm layer nz-2
                const real cqw = R(0.5) * (qt_m + qt_mm);
                const real cq1 = R(1.0) / (R(1.0) + cqw);
                real pressure = ((cq1 * R(2.0)) * rdnw[nz - 1]) * (R(0.0) - pp_m);
                real term = pressure - c1f[nz] * mupc;
                term = term - (cqw * cq1) * (c1f[nz] * mubc + c2f[nz]);
                t = t + G_F * term;
            } else {
                real pressure = ((R(1.0) * R(2.0)) * rdnw[nz - 1]) * (R(0.0) - pp_m);
                real term = pressure - c1f[nz] * mupc;
                t = t + G_F * term;
            }
        }
        if (DIFF) {
            real lap = R(0.0);
            if (interior) {
                real fc = wz[2];
                lap = s.kh * (s.dxi2 * ((wx4 - R(2.0) * fc) + wx2)
                              + s.dyi2 * ((wy4 - R(2.0) * fc) + wy2));
                real fup = (wz[3] - fc) * rdzf_w[k];
                real fdn = (fc - wz[1]) * rdzf_w[k - 1];
                lap = lap + (s.kv * (fup - fdn)) * rdzc_w[k];
            }
            t = t + (c1f[k] * mA + c2f[k]) * lap;
        }
        if (HELD) t = t + gld(d.d6_w, lm + rA, cC);   // held, last
        gst(d.rw_t, lm + rA, cC, t);

        // --- rph_t: WRF rhs_ph ---
        real r = R(0.0);
        if (K != 0) {
            if (!top) {
                real dph_lo = ph_0 - ph_m, dph_hi = ph_p - ph_0;
                if (PREC) {
                    dph_lo = dphb(k - 1) + dph_lo;
                    dph_hi = dphb(k) + dph_hi;
                }
                real wd_lo = ((R(0.5) * (ww_0 + ww_m)) * rdnw[k - 1]) * dph_lo;
                real wd_hi = ((R(0.5) * (ww_p + ww_0)) * rdnw[k]) * dph_hi;
                real omega = fnm[k] * wd_hi + fnp[k] * wd_lo;
                r = r - omega;
            }
            real mass_k = c1f[k] * mA + c2f[k];
            r = r + mass_k * (G_F * wz[2]);
            real qxk = top ? R(2.0) * qx : qx, qyk = top ? R(2.0) * qy : qy;
            real wi, wip, wj, wjp;
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
            real fcx_i = (c1f[k] * mux_i + c2f[k]) * wi;
            real fcx_ip = (c1f[k] * mux_ip + c2f[k]) * wip;
            real fcy_j = (c1f[k] * muy_j + c2f[k]) * wj;
            real fcy_jp = (c1f[k] * muy_jp + c2f[k]) * wjp;
            // the window holds php alone in the precision form
            real p0 = (PREC && B3) ? phb_at(k, rA, cC) + ph_0 : ph_0;
            if (XO && s.h5) {
                // h_sca_adv_order = 5 (reference dycore.cu:668-748): the
                // face-flux sum times (45 d1 - 9 d2 + d3) / 60 of the
                // centred differences d_o = ph(+o) - ph(-o)
                auto cen = [&](real d1, real d2, real d3) {
                    return div_const((R(45.0) * d1 + R(-9.0) * d2)
                                         + R(1.0) * d3,
                                     1.0 / 60.0, R(60.0));
                };
                const real dx5 = cen(phh_at(k, rA, cx[4]) - phh_at(k, rA, cx[2]),
                                     phh_at(k, rA, cx[5]) - phh_at(k, rA, cx[1]),
                                     phh_at(k, rA, cx[6]) - phh_at(k, rA, cx[0]));
                const real dy5 = cen(phh_at(k, yo[4], cC) - phh_at(k, yo[2], cC),
                                     phh_at(k, yo[5], cC) - phh_at(k, yo[1], cC),
                                     phh_at(k, yo[6], cC) - phh_at(k, yo[0], cC));
                const real hx = (qxk * (fcx_i + fcx_ip)) * dx5;
                const real hy = (qyk * (fcy_j + fcy_jp)) * dy5;
                r = r - (hx + hy);
            } else {
                real fx = fcx_i * (p0 - phh_at(k, rA, cx[2]));
                real fxp = fcx_ip * (phh_at(k, rA, cx[4]) - p0);
                r = r - qxk * (fxp + fx);
                real fy = fcy_j * (p0 - phh_at(k, rS, cC));
                real fyp = fcy_jp * (phh_at(k, rN, cC) - p0);
                r = r - qyk * (fyp + fy);
            }
        }
        gst(d.rph_t, lm + rA, cC, r);

        // --- w'' and phi'' initialisation at w level k ---
        gst(d.w_pp, lm + rA, cC,
            (c1f[k] * mt + c2f[k]) * gld(d.w0, lm + rA, cC)
                - (c1f[k] * mA + c2f[k]) * wz[2]);
        gst(d.ph_pp, lm + rA, cC, phpp_0);

        // --- mass level k: rth_t and the theta''/p''/alpha'' init ---
        real phpp_p = R(0.0);
        if (mass) {
            real qxv[7], qyv[7], tpx[7], tpy[7];
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
            real fx0 = flux5(qxv[0], qxv[1], qxv[2], qxv[3], qxv[4], qxv[5], ru_i);
            real fx1 = flux5(qxv[1], qxv[2], qxv[3], qxv[4], qxv[5], qxv[6], ru_ip);
            real fy0 = flux5(qyv[0], qyv[1], qyv[2], qyv[3], qyv[4], qyv[5], rv_j);
            real fy1 = flux5(qyv[1], qyv[2], qyv[3], qyv[4], qyv[5], qyv[6], rv_jp);
            const real fz1 = zf(thz[1], thz[2], thz[3], thz[4], ww_p, k + 1, nz, fnm, fnp);
            real t2 = div3(fx1 - fx0, fy1 - fy0, fz1 - fzt, s.rdx, s.rdy, rdnw[k]);
            fzt = fz1;
            if (XO && MOIST && d.h_diabatic) {
                // the retained microphysics heating, coupled with the
                // stage's column mass (reference dycore.py:572-592)
                t2 = t2 + (c1h[k] * mA + c2h[k]) * gld(d.h_diabatic, lm + rA, cC);
            }
            if (DIFF) {
                real fc = thp_0;
                real lap = s.kh * (s.dxi2 * ((tpx[4] - R(2.0) * fc) + tpx[2])
                                    + s.dyi2 * ((tpy[4] - R(2.0) * fc) + tpy[2]));
                const real gup = (thp_p - fc) * rdzf_h[k < nz - 1 ? k : nz - 2];
                const real gdn = (fc - thp_m) * rdzf_h[k > 0 ? k - 1 : 0];
                real fup = (k < nz - 1) ? gup : R(0.0);
                real fdn = (k > 0) ? gdn : R(0.0);
                lap = lap + (s.kv * (fup - fdn)) * rdzc_h[k];
                t2 = t2 + (c1h[k] * mA + c2h[k]) * lap;
            }
            if (HELD) t2 = t2 + gld(d.d6_t, lm + rA, cC);
            gst(d.rth_t, lm + rA, cC, t2);

            phpp_p = gld(d.php0, lm + st + rA, cC) - gld(d.php, lm + st + rA, cC);
            real cht = c1h[k] * mt + c2h[k];
            real chs = c1h[k] * mA + c2h[k];
            real thbk = thb_at(k, rA, cC);
            const real thp0k = gld(d.thp0, lm + rA, cC);
            real tht = thbk + thp0k;
            real ths = thbk + thp_0;
            real thpp;
            if (PREC)
                thpp = (c1h[k] * mupp) * thbk + (cht * thp0k - chs * thp_0);
            else
                thpp = cht * tht - chs * ths;
            gst(d.th_pp, lm + rA, cC, thpp);
            real altk = gld(d.alt, lm + rA, cC);
            real c2a = (GAMMA_F * p_k) / altk;
            real al = -((altk * (c1h[k] * mupp)) + rdnw[k] * (phpp_p - phpp_0)) / cht;
            if (B3) gst(d.al_pp, lm + rA, cC, al);
            real p5 = (altk * (thpp - (c1h[k] * mupp) * ths)) / (cht * ths);
            gst(d.p_pp, lm + rA, cC, c2a * (p5 - al));
        }
        (void)km;

        // shift windows (refills are clamped loads and selects)
        phpp_0 = phpp_p;
        const real th_n = th_at(k + 3 < nz ? k + 3 : nz - 1, rA, cC);
        thz[0] = thz[1]; thz[1] = thz[2]; thz[2] = thz[3]; thz[3] = thz[4];
        thz[4] = (k + 3 < nz) ? th_n : R(0.0);
        const real w_n = gld(d.w, (k + 3 <= nz ? k + 3 : nz) * st + rA, cC);
        wz[0] = wz[1]; wz[1] = wz[2]; wz[2] = wz[3]; wz[3] = wz[4];
        wz[4] = (k + 3 <= nz) ? w_n : R(0.0);
        const real ww_n = gld(d.ww, (k + 2 <= nz ? k + 2 : nz) * st + rA, cC);
        ww_m = ww_0; ww_0 = ww_p;
        ww_p = (k + 2 <= nz) ? ww_n : R(0.0);
        const real ph_n = phw_at(k + 2 <= nz ? k + 2 : nz, rA, cC);
        ph_m = ph_0; ph_0 = ph_p;
        ph_p = (k + 2 <= nz) ? ph_n : R(0.0);
        pp_m = pp_0;
        if (MOIST) {
            qt_mm = qt_m;
            qt_m = qt_0;
        }
        ru_i_m = ru_i; ru_ip_m = ru_ip; rv_j_m = rv_j; rv_jp_m = rv_jp;
        u_i_mm = u_i_m; u_i_m = u_i; u_ip_mm = u_ip_m; u_ip_m = u_ip;
        v_j_mm = v_j_m; v_j_m = v_j; v_jp_mm = v_jp_m; v_jp_m = v_jp;
        const real thp_n = gld(d.thp, (k + 2 < nz ? k + 2 : nz - 1) * st + rA, cC);
        thp_m = thp_0; thp_0 = thp_p;
        thp_p = (k + 2 < nz) ? thp_n : R(0.0);
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
    real pA = R(0.0), oA = R(0.0), pW = R(0.0), oW = R(0.0), pS = R(0.0), oS = R(0.0);
    real phA = R(0.0), phW = R(0.0), phS = R(0.0);   // ph'' at w level k+1
    real hpA = R(0.0), hpW = R(0.0), hpS = R(0.0);   // php at w level k+1
    real gbA = R(0.0), gbW = R(0.0), gbS = R(0.0);   // phb at w level k+1 (terrain)
    real altA = R(0.0), altW = R(0.0), altS = R(0.0);
    real aA = R(0.0), aW = R(0.0), aS = R(0.0);      // al'' (terrain)
    real bA = R(0.0), bW = R(0.0), bS = R(0.0);      // pb (terrain)
    real u = R(0.0), ru = R(0.0), v = R(0.0), rv = R(0.0);
};

// u'' at face (j, iu) and/or v'' at face (jv, i) through all levels
// (reference acoustic.cu:244-314).  A is the mass column east of the u face
// and north of the v face, W its west and S its south neighbour.  The
// damped pressure p'' + smdiv (p'' - p''_old) rides in 4-level windows for
// the surface and lid extrapolations of dpn.
template <bool B3, bool USEP, bool DU, bool DV, bool SUM = false,
          bool XO = true>
__device__ __forceinline__ void uv_march(const Dom& d, const Sc& s, int rA,
                                         int rS, int rU, unsigned bA,
                                         unsigned bW, unsigned bS, unsigned bu,
                                         unsigned bv)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const real smdiv = s.smdiv, dtau = s.dtau;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    kreal* rdnw = kview(d.rdnw);
    auto opaque_all = [&]() {
        colb_opaque(bA); colb_opaque(bW); colb_opaque(bS);
        colb_opaque(bu); colb_opaque(bv);
    };
    auto pd = [&](real a, real o) {
        real b = USEP ? o : a;
        return a + smdiv * (a - b);
    };
    const real mA = gld(d.mub2d, rA, bA) + gld(d.mup, rA, bA);
    const real muf_u = DU ? R(0.5) * ((gld(d.mub2d, rA, bW) + gld(d.mup, rA, bW)) + mA)
                           : R(0.0);
    const real muf_v = DV ? R(0.5) * ((gld(d.mub2d, rS, bS) + gld(d.mup, rS, bS)) + mA)
                           : R(0.0);
    const real mppA = gld(d.mu_pp, rA, bA);
    const real dmu_u = DU ? gld(d.mu_pp, rA, bW) + mppA : R(0.0);
    const real dmu_v = DV ? gld(d.mu_pp, rS, bS) + mppA : R(0.0);

    // windows S(k-2), S(k-1), S(k), S(k+1) of the damped pressure
    opaque_all();
    real pa[4] = {R(0.0), R(0.0), R(0.0), R(0.0)}, pw[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
    real ps[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
    pa[2] = pd(gld(d.p_pp, 0 + rA, bA), gld(d.p_pp_old, 0 + rA, bA));
    pa[3] = pd(gld(d.p_pp, st + rA, bA), gld(d.p_pp_old, st + rA, bA));
    const real pa2 = pd(gld(d.p_pp, 2 * st + rA, bA), gld(d.p_pp_old, 2 * st + rA, bA));
    real dnu_0 = R(0.0), dnv_0 = R(0.0);
    if (DU) {
        pw[2] = pd(gld(d.p_pp, 0 + rA, bW), gld(d.p_pp_old, 0 + rA, bW));
        pw[3] = pd(gld(d.p_pp, st + rA, bW), gld(d.p_pp_old, st + rA, bW));
        const real pw2 = pd(gld(d.p_pp, 2 * st + rA, bW), gld(d.p_pp_old, 2 * st + rA, bW));
        real S0 = pa[2] + pw[2], S1 = pa[3] + pw[3], S2 = pa2 + pw2;
        dnu_0 = R(0.5) * ((d.cf1 * S0 + d.cf2 * S1) + d.cf3 * S2);
    }
    if (DV) {
        ps[2] = pd(gld(d.p_pp, 0 + rS, bS), gld(d.p_pp_old, 0 + rS, bS));
        ps[3] = pd(gld(d.p_pp, st + rS, bS), gld(d.p_pp_old, st + rS, bS));
        const real ps2 = pd(gld(d.p_pp, 2 * st + rS, bS), gld(d.p_pp_old, 2 * st + rS, bS));
        real S0 = pa[2] + ps[2], S1 = pa[3] + ps[3], S2 = pa2 + ps2;
        dnv_0 = R(0.5) * ((d.cf1 * S0 + d.cf2 * S1) + d.cf3 * S2);
    }
    real phA0 = gld(d.ph_pp, 0 + rA, bA), hpA0 = gld(d.php, 0 + rA, bA);
    real phW0 = DU ? gld(d.ph_pp, 0 + rA, bW) : R(0.0), hpW0 = DU ? gld(d.php, 0 + rA, bW) : R(0.0);
    real phS0 = DV ? gld(d.ph_pp, 0 + rS, bS) : R(0.0), hpS0 = DV ? gld(d.php, 0 + rS, bS) : R(0.0);
    real gbA0 = B3 ? gld(d.phb, 0 + rA, bA) : R(0.0);
    real gbW0 = (B3 && DU) ? gld(d.phb, 0 + rA, bW) : R(0.0);
    real gbS0 = (B3 && DV) ? gld(d.phb, 0 + rS, bS) : R(0.0);

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
        const real c1 = c1h[k], c2 = c2h[k];
        const real phA1 = L.phA, hpA1 = L.hpA, gbA1 = L.gbA;
        const real altA = L.altA;
        auto dn_top = [&](real Sk, real Sm1, real Sm2) {
            return s.top_lid ? R(0.5) * ((d.cf1 * Sk + d.cf2 * Sm1) + d.cf3 * Sm2)
                             : R(0.0);
        };
        if (DU) {
            real Sk = pa[2] + pw[2];
            real dn_up;
            if (LAST) {
                dn_up = dn_top(Sk, pa[1] + pw[1], pa[0] + pw[0]);
            } else {
                real Sp = pa[3] + pw[3];
                dn_up = R(0.5) * (fnm[k + 1] * Sp + fnp[k + 1] * Sk);
            }
            real dph = (phA1 - L.phW) + (phA0 - phW0);
            real dpe = pa[2] - pw[2];
            real brk = dph + (altA + L.altW) * dpe;
            if (B3) brk = brk + (L.aA + L.aW) * (L.bA - L.bW);
            real dpxy = ((R(0.5) * s.rdx) * (c1 * muf_u + c2)) * brk;
            real phsA = hpA0 + hpA1, phsB = hpW0 + L.hpW;
            real dphp = R(0.5) * (phsA - phsB);
            if (B3) dphp = dphp + R(0.5) * ((gbA0 + gbA1) - (gbW0 + L.gbW));
            dpxy = dpxy + (s.rdx * dphp) * (rdnw[k] * (dn_up - dnu_0)
                                            - (R(0.5) * c1) * dmu_u);
            if (XO && s.cq) dpxy = gld(d.cqu, k * stu + rU, bu) * dpxy;   // moist_cq
            const real un = L.u + dtau * (L.ru - dpxy);
            gst(d.u_pp, k * stu + rU, bu, un);
            if (SUM) {
                const real acc = (s.sum == 1) ? R(0.0)
                                              : gld(d.ru_m, k * stu + rU, bu);
                gst(d.ru_m, k * stu + rU, bu, acc + un);
            }
            dnu_0 = dn_up;
            phW0 = L.phW; hpW0 = L.hpW; gbW0 = L.gbW;
        }
        if (DV) {
            real Sk = pa[2] + ps[2];
            real dn_up;
            if (LAST) {
                dn_up = dn_top(Sk, pa[1] + ps[1], pa[0] + ps[0]);
            } else {
                real Sp = pa[3] + ps[3];
                dn_up = R(0.5) * (fnm[k + 1] * Sp + fnp[k + 1] * Sk);
            }
            real dph = (phA1 - L.phS) + (phA0 - phS0);
            real dpe = pa[2] - ps[2];
            real brk = dph + (altA + L.altS) * dpe;
            if (B3) brk = brk + (L.aA + L.aS) * (L.bA - L.bS);
            real dpxy = ((R(0.5) * s.rdy) * (c1 * muf_v + c2)) * brk;
            real phsA = hpA0 + hpA1, phsB = hpS0 + L.hpS;
            real dphp = R(0.5) * (phsA - phsB);
            if (B3) dphp = dphp + R(0.5) * ((gbA0 + gbA1) - (gbS0 + L.gbS));
            dpxy = dpxy + (s.rdy * dphp) * (rdnw[k] * (dn_up - dnv_0)
                                            - (R(0.5) * c1) * dmu_v);
            if (XO && s.cq) dpxy = gld(d.cqv, k * stv + rA, bv) * dpxy;   // moist_cq
            const real vn = L.v + dtau * (L.rv - dpxy);
            gst(d.v_pp, k * stv + rA, bv, vn);
            if (SUM) {
                const real acc = (s.sum == 1) ? R(0.0)
                                              : gld(d.rv_m, k * stv + rA, bv);
                gst(d.rv_m, k * stv + rA, bv, acc + vn);
            }
            dnv_0 = dn_up;
            phS0 = L.phS; hpS0 = L.hpS; gbS0 = L.gbS;
        }
        phA0 = phA1; hpA0 = hpA1; gbA0 = gbA1;
        // advance the damped-pressure windows to k+1 (level k+2 enters)
        const bool in = k + 2 < nz;
        pa[0] = pa[1]; pa[1] = pa[2]; pa[2] = pa[3];
        pa[3] = in ? pd(L.pA, L.oA) : R(0.0);
        if (DU) {
            pw[0] = pw[1]; pw[1] = pw[2]; pw[2] = pw[3];
            pw[3] = in ? pd(L.pW, L.oW) : R(0.0);
        }
        if (DV) {
            ps[0] = ps[1]; ps[1] = ps[2]; ps[2] = ps[3];
            ps[3] = in ? pd(L.pS, L.oS) : R(0.0);
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
// Batched (Many): each member's logical range holds its edge workgroups
// first; ne_arg is unused.
template <class SRC, bool B3, bool USEP, bool SUM = false, bool XO = true>
__global__ __launch_bounds__(256) void k_uv(typename SRC::Arg a, Sc s, int ne_arg)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const int ny = d.ny, nx = d.nx, pu = nx + 1;
    const int ne = SRC::many ? (ny + nx + 255) / 256 : ne_arg;
    if (g < ne) {
        const int t = g * 256 + threadIdx.y * 64 + threadIdx.x;
        if (t < ny) {
            const int j = t;
            uv_march<B3, USEP, true, false, SUM, XO>(
                d, s, 0, 0, 0, RB * (j * nx), RB * (j * nx + nx - 1), 0u,
                RB * (j * pu + nx), 0u);
        } else if (t < ny + nx) {
            const int i = t - ny;
            uv_march<B3, USEP, false, true, SUM, XO>(
                d, s, 0, 0, 0, RB * i, 0u, RB * ((ny - 1) * nx + i), 0u,
                RB * (ny * nx + i));
        }
        return;
    }
    const RowSeg rs = SRC::many ? row_seg_at(nx, ny, g - ne)
                              : row_seg_from(nx, ny, ne);
    if (!rs.live) return;
    const int i = rs.i, j = rs.j;
    const int jm = (j == 0) ? ny - 1 : j - 1;
    const unsigned bc = RB * i, bw = RB * ((i == 0) ? nx - 1 : i - 1);
    uv_march<B3, USEP, true, true, SUM, XO>(d, s, j * nx, jm * nx, j * pu, bc, bw, bc,
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
    real ui = R(0.0), uip = R(0.0);   // u'' at faces i, i+1
    real vj = R(0.0), vjp = R(0.0);   // v'' at faces j, j+1
    real te = R(0.0), tw = R(0.0), tn = R(0.0), ts = R(0.0);  // thp, 4 neighbours
    real be = R(0.0), bw = R(0.0), bn = R(0.0), bs = R(0.0);  // thb there (terrain)
    real thu = R(0.0), bhu = R(0.0);  // thp, thb at (m+1, c)
    real thpp = R(0.0), rth = R(0.0); // theta'' and its tendency
    real phu = R(0.0), ghu = R(0.0);  // php, phb at w level m+1
    real alt = R(0.0), p = R(0.0);    // stage reference alt, p
    real phpp = R(0.0), wpp = R(0.0), rph = R(0.0);  // at w level m
    real rw = R(0.0);               // w tendency at w level m-1
};

// Inputs of downward w level k (and mass level k when k < nz).
struct ColDn {
    real wpp = R(0.0), gam = R(0.0), w = R(0.0), php = R(0.0), phb = R(0.0);
    real phpp = R(0.0), thp = R(0.0), thb = R(0.0), alt = R(0.0), p = R(0.0);
    real thpp = R(0.0);
    double wpp64 = 0.0, gam64 = 0.0;      // precision form
};


// B3: terrain (3-D) base state; DAMP: damp_opt = 3 w damper; AL: store
// alpha''; WW: store Omega'' (standalone substep API).  nz >= 3.
// Batched (Many): the new p'' goes to each member's p_pp_old.
// Rows r >= nz + 1 - s.park keep gamma and the eliminated w'' in LDS
// (2 s.park values per column, dynamic shared memory) instead of parking
// them in HBM between the sweeps; rows below go to d.gam / d.w_pp, gamma
// only when the stage has not stored it yet (s.gam_ready = 0).
template <class SRC, bool B3, bool DAMP, bool AL, bool WW, bool PREC,
          bool XO = true>
__global__ __launch_bounds__(256) void k_column(typename SRC::Arg a, Sc s,
                                                real* RS pout_arg)
{
    extern __shared__ real lpark[];
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    real* RS pout = SRC::many ? d.p_pp_old : pout_arg;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const int c = (SRC::many ? g : xcd_block()) * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
    const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
    const real eps = s.epssm, dtau = s.dtau;
    const real one_p = R(1.0) + eps, one_m = R(1.0) - eps;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* c1f = kview(d.c1f);
    kreal* c2f = kview(d.c2f);
    kreal* dnw = kview(d.dnw);
    kreal* rdnw = kview(d.rdnw);
    kreal* rdn = kview(d.rdn);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    // per-thread column byte offsets
    unsigned bc = RB * c;
    unsigned bE = RB * (j * nx + ip), bW = RB * (j * nx + im);
    unsigned bN = RB * (jp * nx + i), bS = RB * (jm * nx + i);
    unsigned bu = RB * (j * (nx + 1) + i), bu1 = bu + RB;
    unsigned bvn = RB * (c + nx);
    auto opaque_all = [&]() {
        colb_opaque(bc); colb_opaque(bE); colb_opaque(bW);
        colb_opaque(bN); colb_opaque(bS); colb_opaque(bu);
        colb_opaque(bu1); colb_opaque(bvn);
    };

    const real mut = d.mub2d[c] + d.mup[c];
    const real mupp_old = d.mu_pp[c];
    real dmdt_pp = R(0.0);
    Kahan kdm;
#pragma unroll 5
    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const real a = gld(d.u_pp, k * stu, bu1), b = gld(d.u_pp, k * stu, bu);
        const real e = gld(d.v_pp, k * stv, bvn), f = gld(d.v_pp, k * stv, bc);
        const real x = dnw[k] * (s.rdx * (a - b) + s.rdy * (e - f));
        if (PREC) kdm.add(x);
        else dmdt_pp = dmdt_pp + x;
    }
    if (PREC) dmdt_pp = kdm.s;
    // the tridiagonal's type: FP64 in the precision form
    typedef typename std::conditional<PREC, double, real>::type tri;
    const int stt = (nz + 1) * st;             // tri64: gamma, then w''

    const real rmu = d.rmu_t[c];
    const real mupp = mupp_old + dtau * ((dmdt_pp + d.dmdt_rf[c]) + rmu);
    d.mu_pp[c] = mupp;
    const real muts = mut + mupp;
    const real muave = R(0.5) * (one_p * mupp + one_m * mupp_old);
    real cof = R(0.5) * dtau * G_F * one_p;
    cof = cof * cof;
    auto cfm_at = [&](int kk) { return c1f[kk] * mut + c2f[kk]; };
    // first row parked in LDS; slot (r - r0) holds gamma at 2 (r - r0),
    // the eliminated w'' at 2 (r - r0) + 1, each a row of blockDim.x
    const int r0 = PREC ? nz + 2 : nz + 1 - s.park;
    const int lt = (int)blockDim.x, lx = (int)threadIdx.x;
    kreal* thb1 = kview(d.thb);     // 1-D base profiles (flat base)
    kreal* phb1 = kview(d.phb);
    auto thbv = [&](real loaded, int kk) { return B3 ? loaded : thb1[kk]; };
    auto phbv = [&](real loaded, int kk) { return B3 ? loaded : phb1[kk]; };

    // kinematic lower boundary on w'' from the updated u''/v''
    real w0;
    {
        opaque_all();
        auto lw = [&](const real* a, unsigned colb, int sl) {
            return (d.cf1 * gld(a, 0, colb) + d.cf2 * gld(a, sl, colb))
                   + d.cf3 * gld(a, 2 * sl, colb);
        };
        real ue = lw(d.u_pp, bu1, stu), uw = lw(d.u_pp, bu, stu);
        real vn = lw(d.v_pp, bvn, stv), vs = lw(d.v_pp, bc, stv);
        const real* RS ht = d.ht;
        real h = ht[c];
        w0 = (R(0.5) * s.rdy) * ((ht[jp * nx + i] - h) * vn + (h - ht[jm * nx + i]) * vs)
             + (R(0.5) * s.rdx) * ((ht[j * nx + ip] - h) * ue + (h - ht[j * nx + im]) * uw);
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
    // Omega'' at w level q: stored (standalone substep), or summed into
    // ww_m (the step's fused sumflux)
    auto put_ww = [&](int q, real x) {
        if (!WW) return;
        if (s.sum == 0) {
            gst(d.ww_pp, q * st, bc, x);
        } else {
            const real acc = (s.sum == 1) ? R(0.0) : gld(d.ww_m, q * st, bc);
            gst(d.ww_m, q * st, bc, acc + x);
        }
    };
    put_ww(0, R(0.0));
    real ww_k = R(0.0), wwk = R(0.0), wdtn_k = R(0.0);
    Kahan kww;
    real th_m = (B3 ? gld(d.thb, 0, bc) : thb1[0]) + gld(d.thp, 0, bc);
    real ph_m = (B3 ? gld(d.phb, 0, bc) : phb1[0]) + gld(d.php, 0, bc);
    real wd_prev = R(0.0);
    real rhs_a = R(0.0), rhs_b = R(0.0);
    real pho_a = R(0.0), pho_b = gld(d.ph_pp, 0, bc);
    real t2_a = R(0.0), t2_b = R(0.0);
    real c2a_a = R(0.0), c2a_b = R(0.0);
    real alt_a = R(0.0), alt_b = R(0.0);
    real chm_a = R(0.0), chm_b = R(0.0);
    real wold_b = R(0.0);
    tri wf_prev = (tri)w0;
    tri gam_prev = (tri)R(0.0);
    const real ghalf_eps = (R(0.5) * G_F) * one_m;
    const real c_force = (R(0.5) * dtau) * G_F;
    const real c_buoy = dtau * G_F;

    auto force_row = [&](int r, real rwr, real rhs_dn, real rhs_c,
                         real rhs_up, real ph_dn, real ph_c, real ph_up,
                         real wold, real t2dn, real t2c, real c2dn,
                         real c2c, real altdn, real altc, real chmdn,
                         real chmc) {
        real dph_up = one_p * (rhs_up - rhs_c) + one_m * (ph_up - ph_c);
        real dph_dn = one_p * (rhs_c - rhs_dn) + one_m * (ph_c - ph_dn);
        real pg = (c_force * rdn[r])
                  * ((c2c * rdnw[r]) / chmc * dph_up
                     - (c2dn * rdnw[r - 1]) / chmdn * dph_dn);
        if (XO && s.cq) pg = gld(d.cqw, r * st, bc) * pg;            // moist_cq
        real t1 = dtau * rwr + pg;
        real t3 = c_buoy * (rdn[r] * ((c2c * altc) * t2c - (c2dn * altdn) * t2dn)
                             - c1f[r] * muave);
        return wold + (t1 + t3);
    };
    // LU factors of row r (1 <= r <= nz-1) from the windowed c2a/chm
    auto lu_row = [&](int r, real c2dn, real c2c, real chmdn, real chmc,
                      tri& alpha_r, bool r_ge2) {
        // moist_cq: the row's coefficient takes cqw(r) first
        const tri tcof = (XO && s.cq) ? (tri)gld(d.cqw, r * st, bc) * (tri)cof
                              : (tri)cof;
        const tri rdn_r = rdn[r], rdnw_r = rdnw[r];
        const tri rdnw_m = rdnw[r - 1];
        tri a_r = (tri)R(0.0);
        if (r_ge2)
            a_r = -tcof * rdn_r * rdnw_m * (tri)c2dn
                  / ((tri)chmdn * (tri)cfm_at(r - 1));
        tri b = (tri)R(1.0) + tcof * rdn_r
                                  * (rdnw_r * (tri)c2c / ((tri)chmc * (tri)cfm_at(r))
                                     + rdnw_m * (tri)c2dn
                                           / ((tri)chmdn * (tri)cfm_at(r)));
        tri cc = -tcof * rdn_r * rdnw_r * (tri)c2c
                 / ((tri)chmc * (tri)cfm_at(r + 1));
        alpha_r = (tri)R(1.0) / (b - a_r * gam_prev);
        gam_prev = cc * alpha_r;
        if (PREC) {
            if (!s.gam_ready) d.tri64[r * st + c] = (double)gam_prev;
        } else if (r >= r0) {
            lpark[(2 * (r - r0)) * lt + lx] = (real)gam_prev;
        } else if (!s.gam_ready) {
            gst(d.gam, r * st, bc, (real)gam_prev);
        }
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
        real dvp = s.rdx * (L.uip - L.ui) + s.rdy * (L.vjp - L.vj);
        real ww_kp1;
        if (has_up) {
            const real inc = dnw[m] * (c1h[m] * (dmdt_pp + rmu) + dvp);
            if (PREC) {
                kww.add(-inc);
                wwk = kww.s;
            } else {
                wwk = wwk - inc;
            }
            ww_kp1 = wwk;
        } else {
            ww_kp1 = R(0.0);
        }
        put_ww(m + 1, ww_kp1);
        real th_up = has_up ? thbv(L.bhu, m + 1) + L.thu : R(0.0);
        real wdtn_kp1 = has_up
            ? ww_kp1 * (fnm[m + 1] * th_up + fnp[m + 1] * th_m) : R(0.0);
        real t_c = th_m;
        real hx = (R(0.5) * s.rdx) * (L.uip * (thbv(L.be, m) + L.te + t_c)
                                     - L.ui * (t_c + (thbv(L.bw, m) + L.tw)));
        real hy = (R(0.5) * s.rdy) * (L.vjp * (thbv(L.bn, m) + L.tn + t_c)
                                     - L.vj * (t_c + (thbv(L.bs, m) + L.ts)));
        real tho = L.thpp;
        real thn = tho + dtau * (L.rth - ((hx + hy) + rdnw[m] * (wdtn_kp1 - wdtn_k)));
        gst(d.th_pp, ms, bc, thn);
        real chs = c1h[m] * muts + c2h[m];
        real t2 = (R(0.5) * (one_p * thn + one_m * tho)) / (chs * t_c);
        real ph_up = phbv(L.ghu, m + 1) + L.phu;
        real wd = ((R(0.5) * (ww_kp1 + ww_k)) * rdnw[m]) * (ph_up - ph_m);
        real altm = L.alt;
        real c2a = GAMMA_F * L.p / altm;
        real chm = c1h[m] * mut + c2h[m];
        real rhs_c = R(0.0), pho_c = pho_b, wold_c = R(0.0);
        if (has_rhs) {
            pho_c = L.phpp;
            wold_c = L.wpp;
            real r = dtau * (L.rph + ghalf_eps * wold_c);
            r = r - dtau * (fnm[m] * wd + fnp[m] * wd_prev);
            rhs_c = pho_c + r / cfm_at(m);
            gst(d.ph_pp, ms, bc, rhs_c);
        }
        if (has_row) {
            const int r = m - 1;
            tri alpha_r;
            const bool r_ge2 = K == LV_MID ? true
                               : K == LV_THIRD ? false : r >= 2;
            tri a_r = lu_row(r, c2a_a, c2a_b, chm_a, chm_b, alpha_r, r_ge2);
            real wr = force_row(r, L.rw, rhs_a, rhs_b, rhs_c, pho_a, pho_b,
                                 pho_c, wold_b, t2_a, t2_b, c2a_a, c2a_b,
                                 alt_a, alt_b, chm_a, chm_b);
            tri wf = ((tri)wr - a_r * wf_prev) * alpha_r;
            if (PREC) d.tri64[stt + r * st + c] = (double)wf;
            else if (r >= r0) lpark[(2 * (r - r0) + 1) * lt + lx] = (real)wf;
            else gst(d.w_pp, r * st, bc, (real)wf);
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
        const real pho_c = nxt.phpp;
        const real wold_c = nxt.wpp;
        real r = dtau * (nxt.rph + ghalf_eps * wold_c);
        real rhs_c = pho_c + r / cfm_at(nz);
        if (s.top_lid) rhs_c = R(0.0);
        gst(d.ph_pp, f, bc, rhs_c);
        const int rr = nz - 1;
        tri alpha_r;
        tri a_r = lu_row(rr, c2a_a, c2a_b, chm_a, chm_b, alpha_r, rr >= 2);
        real wr = force_row(rr, nxt.rw, rhs_a, rhs_b, rhs_c, pho_a, pho_b,
                             pho_c, wold_b, t2_a, t2_b, c2a_a, c2a_b, alt_a,
                             alt_b, chm_a, chm_b);
        tri wfr = ((tri)wr - a_r * wf_prev) * alpha_r;
        if (PREC) d.tri64[stt + rr * st + c] = (double)wfr;
        else if (rr >= r0) lpark[(2 * (rr - r0) + 1) * lt + lx] = (real)wfr;
        else gst(d.w_pp, rr * st, bc, (real)wfr);
        tri a_top = (tri)R(0.0);
        real wtop;
        if (s.top_lid) {
            wtop = R(0.0);
        } else {
            a_top = -(tri)R(2.0) * (tri)cof * (tri)rdnw[nz - 1]
                    * (tri)rdnw[nz - 1] * (tri)c2a_b
                    / ((tri)chm_b * (tri)cfm_at(nz - 1));
            real dph_dn = one_p * (rhs_c - rhs_b) + one_m * (pho_c - pho_b);
            real bterm = dtau * G_F * c2a_b * rdnw[nz - 1] * rdnw[nz - 1]
                          / chm_b * dph_dn;
            real cterm = dtau * G_F * (R(2.0) * rdnw[nz - 1] * c2a_b * alt_b * t2_b
                                        + c1f[nz] * muave);
            wtop = wold_c + ((dtau * d.rw_t[f + c] - bterm) - cterm);
        }
        tri b_top = (tri)R(1.0) + (tri)R(2.0) * (tri)cof * (tri)rdnw[nz - 1]
                                    * (tri)rdnw[nz - 1] * (tri)c2a_b
                                    / ((tri)chm_b * (tri)cfm_at(nz));
        tri alpha_top = (tri)R(1.0) / (b_top - a_top * gam_prev);
        const tri wtf = ((tri)wtop - a_top * wfr) * alpha_top;
        if (PREC) d.tri64[stt + f + c] = (double)wtf;
        else if (nz >= r0) lpark[(2 * (nz - r0) + 1) * lt + lx] = (real)wtf;
        else gst(d.w_pp, f, bc, (real)wtf);
    }

    // downward sweep: back substitution, damper, phi'' update, EOS
    real htop = R(0.0), hbot = R(0.0);
    if (DAMP) {
        htop = ((B3 ? d.phb[nz * st + c] : phb1[nz]) + d.php[nz * st + c]) / G_F;
        hbot = htop - s.zdamp;
    }
    const real c_ph = ((R(0.5) * dtau) * G_F) * one_p;
    auto load_dn = [&](ColDn& L, int k) {
        const unsigned ks = k * st;
        const unsigned km = (k < nz ? k : nz - 1) * st;
        if (PREC) {
            L.wpp64 = d.tri64[stt + ks + c];
            L.gam64 = d.tri64[ks + c];
        } else if (k >= r0) {
            L.wpp = lpark[(2 * (k - r0) + 1) * lt + lx];
            L.gam = k < nz ? lpark[(2 * (k - r0)) * lt + lx] : R(0.0);
        } else {
            L.wpp = gld(d.w_pp, ks, bc);
            L.gam = gld(d.gam, ks, bc);
        }
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
    real wb_next = R(0.0), phn_up = R(0.0);
    double wb_next64 = 0.0;
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
        real phn;
        if (K != LV_LAST) {
            real wfin;
            if (PREC) {
                double wb = L.wpp64;
                if (K != LV_FIRST) wb = wb - L.gam64 * wb_next64;
                wb_next64 = wb;
                wfin = (real)wb;
            } else {
                real wb = L.wpp;
                if (K != LV_FIRST) wb = wb - L.gam * wb_next;
                wb_next = wb;
                wfin = wb;
            }
            const real wb = wfin;
            if (DAMP) {
                real hk = (phbv(L.phb, k) + L.php) / G_F;
                if (hk >= hbot) {
                    real arg = (R(1.5707963) * (hk - hbot)) / s.zdamp;
                    real sn = (real)sin((double)arg);
                    real dw = (s.dampmag * sn) * sn;
                    wfin = (wb - (dw * cfm_at(k)) * L.w) / (R(1.0) + dw);
                }
            }
            gst(d.w_pp, f, bc, wfin);
            real cfs = c1f[k] * muts + c2f[k];
            phn = L.phpp + (c_ph * wfin) / cfs;
            gst(d.ph_pp, f, bc, phn);
        } else {
            phn = L.phpp;
        }
        if (K != LV_FIRST) {
            real chm2 = c1h[k] * muts + c2h[k];
            real th_ref = thbv(L.thb, k) + L.thp;
            real altk = L.alt;
            real c2a = GAMMA_F * L.p / altk;
            real al = -((altk * (c1h[k] * mupp)) + rdnw[k] * (phn_up - phn)) / chm2;
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
__device__ __forceinline__ void finish_face(const Dom& d, real* RS out,
                                            const real* RS x,
                                            const real* RS xpp, int f,
                                            real msf, real mnf)
{
    const int nz = d.nz, ny = d.ny, nx = d.nx;
    const int sf = STRIDE_KIND == 0 ? ny * (nx + 1) : (ny + 1) * nx;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    const unsigned bf = RB * f;
    // a lone column per thread: unrolled, with no per-level opaque offset,
    // so several levels' loads are in flight at once
#pragma unroll 8
    for (int k = 0; k < nz; ++k) {
        const real c1 = c1h[k], c2 = c2h[k];
        gst(out, k * sf, bf, ((c1 * msf + c2) * gld(x, k * sf, bf)
                              + gld(xpp, k * sf, bf)) / (c1 * mnf + c2));
    }
}

struct FinIn {
    real u = R(0.0), upp = R(0.0), v = R(0.0), vpp = R(0.0);
    real w = R(0.0), wpp = R(0.0), php = R(0.0), phpp = R(0.0);
    real thp = R(0.0), thpp = R(0.0);
    real thb = R(0.0), phb = R(0.0), alb = R(0.0);   // terrain base
    real pb = R(0.0);                                // terrain, PREC
};

// Periodic duplicate faces of the finish: u at i = nx (t < ny, row t),
// v at j = ny (ny <= t < ny + nx, column t - ny).  They read their own
// face's inputs only, so they run beside the interior columns.
__device__ __forceinline__ void finish_edges(const Dom& d, const FinishDst& o,
                                             int t)
{
    const int ny = d.ny, nx = d.nx;
    const real* RS mub = d.mub2d;
    const real* RS mup = d.mup;
    const real* RS mu_pp = d.mu_pp;
    if (t < ny) {
        const int j = t, cA = j * nx, cB = j * nx + nx - 1;
        const real msA = mub[cA] + mup[cA], mnA = msA + mu_pp[cA];
        const real msB = mub[cB] + mup[cB], mnB = msB + mu_pp[cB];
        finish_face<0>(d, o.u, d.u, d.u_pp, j * (nx + 1) + nx,
                       R(0.5) * (msA + msB), R(0.5) * (mnA + mnB));
    } else if (t < ny + nx) {
        const int i = t - ny, cA = i, cB = (ny - 1) * nx + i;
        const real msA = mub[cA] + mup[cA], mnA = msA + mu_pp[cA];
        const real msB = mub[cB] + mup[cB], mnB = msB + mu_pp[cB];
        finish_face<1>(d, o.v, d.v, d.v_pp, ny * nx + i,
                       R(0.5) * (msA + msB), R(0.5) * (mnA + mnB));
    }
}

// The batched launch's duplicate faces (one domain's run inside k_finish).
template <class SRC>
__global__ __launch_bounds__(256) void k_finish_edge(typename SRC::Arg a,
                                                     FinishDst o_arg)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const FinishDst& o = fin_of<SRC>(a, m, o_arg);
    finish_edges(d, o, g * blockDim.x + threadIdx.x);
}

// Small-step finish (reference dycore.cu:1087-1198) of the mass column
// (j, i), its u face (j, i) and v face (j, i), then the EOS of the new
// state (diagnostics.cu:37-89).  PREC (precision form): the new theta' is
// (C(mu_s) thp + theta'' - (c1h mu'') thb) / C(mu_n), free of the base
// theta's cancellation, and the EOS is eos_pert.
// hdt > 0 (the final stage of a step with retained microphysics heating,
// d.h_diabatic set): the theta fold removes hdt C(mu_s) h_diabatic, what
// the substeps integrated from the stage tendency's heating term
// (reference dycore.cu:1186-1194, module_small_step_em.F:408-426).
template <class SRC, bool B3, bool EOS, bool PREC, bool XO = true>
__global__ __launch_bounds__(256) void k_finish(typename SRC::Arg a,
                                                FinishDst o_arg, real hdt,
                                                int nmain)
{
    int m, g;
    if (!SRC::locate(a, m, g)) return;
    const Dom& d = SRC::dom(a, m);
    const FinishDst& o = fin_of<SRC>(a, m, o_arg);
    if (!SRC::many && (int)blockIdx.x >= nmain) {
        // the launch's last workgroups: the periodic duplicate faces
        finish_edges(d, o, ((int)blockIdx.x - nmain) * blockDim.x + threadIdx.x);
        return;
    }
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int stu = ny * (nx + 1), stv = (ny + 1) * nx;
    const int c = (SRC::many ? g : xcd_block_n(nmain)) * blockDim.x + threadIdx.x;
    if (c >= st) return;
    const int j = c / nx, i = c - j * nx;
    const int im = (i == 0) ? nx - 1 : i - 1;
    const int jm = (j == 0) ? ny - 1 : j - 1;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* c1f = kview(d.c1f);
    kreal* c2f = kview(d.c2f);
    kreal* rdnw = kview(d.rdnw);
    kreal* thb1 = kview(d.thb);
    kreal* phb1 = kview(d.phb);
    kreal* alb1 = kview(d.alb);
    kreal* pb1 = kview(d.pb);
    const real* RS mub = d.mub2d;
    const real* RS mup = d.mup;
    const real* RS mu_pp = d.mu_pp;
    const real msA = mub[c] + mup[c], mnA = msA + mu_pp[c];
    const real muppA = mu_pp[c];
    const int cW = j * nx + im, cS = jm * nx + i;
    const real msW = mub[cW] + mup[cW], mnW = msW + mu_pp[cW];
    const real msS = mub[cS] + mup[cS], mnS = msS + mu_pp[cS];
    const real msf_u = R(0.5) * (msA + msW), mnf_u = R(0.5) * (mnA + mnW);
    const real msf_v = R(0.5) * (msA + msS), mnf_v = R(0.5) * (mnA + mnS);
    const real mup_new = mup[c] + mu_pp[c];
    const real mu_eos = mub[c] + mup_new;
    unsigned bc = RB * c, bu = RB * (j * (nx + 1) + i);

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
            if (PREC && EOS) L.pb = gld(d.pb, km * st, bc);
        }
    };
    real php_lo = R(0.0), thp_lo = R(0.0), thb_lo = R(0.0), phb_lo = R(0.0);
    real alb_lo = R(0.0), pb_lo = R(0.0);
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
            const real c1 = c1h[k], c2 = c2h[k];
            gst(o.u, k * stu, bu, ((c1 * msf_u + c2) * L.u + L.upp)
                                      / (c1 * mnf_u + c2));
            gst(o.v, k * stv, bc, ((c1 * msf_v + c2) * L.v + L.vpp)
                                      / (c1 * mnf_v + c2));
        }
        gst(o.w, f, bc, ((c1f[k] * msA + c2f[k]) * L.w + L.wpp)
                            / (c1f[k] * mnA + c2f[k]));
        const real php_new = L.php + L.phpp;
        gst(o.php, f, bc, php_new);
        const real phb_k = B3 ? L.phb : phb1[k];
        if (EOS && K != 0) {
            const unsigned h = f - st;
            if (PREC) {
                const EosPert e = eos_pert(
                    thp_lo, thb_lo, R(0.0), false, alb_lo, pb_lo, mup_new,
                    c1h[k - 1] * mu_eos + c2h[k - 1], c1h[k - 1], rdnw[k - 1],
                    php_new - php_lo);
                gst(d.alt, h, bc, e.alt);
                gst(d.al, h, bc, e.al);
                gst(d.p, h, bc, e.p);
                gst(d.pp, h, bc, e.pp);
            } else {
                real th = thb_lo + thp_lo;
                real dph = (phb_k + php_new) - (phb_lo + php_lo);
                real alt, al;
                eos_alpha<XO>(d, k - 1, mu_eos, dph, alb_lo, c1h[k - 1],
                          c2h[k - 1], rdnw[k - 1], alt, al);
                gst(d.alt, h, bc, alt);
                gst(d.al, h, bc, al);
                gst(d.p, h, bc, P0_F * pow_gamma((RD_F * th) / (P0_F * alt)));
            }
        }
        if (K != 2) {
            const real thbk = B3 ? L.thb : thb1[k];
            real thn;
            if (PREC) {
                real num = ((c1h[k] * msA + c2h[k]) * L.thp + L.thpp)
                           - (c1h[k] * muppA) * thbk;
                thn = num / (c1h[k] * mnA + c2h[k]);
            } else {
                const real cs = c1h[k] * msA + c2h[k];
                real num = cs * (thbk + L.thp) + L.thpp;
                if (XO && hdt > R(0.0))
                    num = num - (hdt * cs) * gld(d.h_diabatic, f, bc);
                thn = num / (c1h[k] * mnA + c2h[k]) - thbk;
            }
            gst(o.thp, f, bc, thn);
            thp_lo = thn;
            thb_lo = thbk;
            alb_lo = B3 ? L.alb : alb1[k];
            if (PREC && EOS) pb_lo = B3 ? L.pb : pb1[k];
        }
        php_lo = php_new;
        phb_lo = phb_k;
    };
    level(0, Kind<0>{});
    for (int k = 1; k < nz; ++k) level(k, Kind<1>{});
    level(nz, Kind<2>{});
    o.mup[c] = mup_new;
}


// ---------------------------------------------------------------------------
// Scalar transport of one RK stage as column marches (reference moist.py:
// 434-698 with the update kernel :353-402 and pd_advection.cu's periodic
// path).  Row-segment waves as in the stencil kernels: each lane marches
// one mass column upward, the y stencil rows are wave-uniform, and the
// vertical flux through a level's upper face is carried to the next level
// as its lower one.
// ---------------------------------------------------------------------------

// WRF flux_upwind (reference pd_advection.cu:101-108).
__device__ __forceinline__ real upwind_flux(real qm1, real q0v, real cr)
{
    return (R(0.5) * fmin(R(1.0), cr + fabs(cr))) * qm1
           + (R(0.5) * fmax(-R(1.0), cr - fabs(cr))) * q0v;
}

// Per-lane view of a mass column for the scalar marches: the 7 x columns
// (periodic), the 7 y rows (wave-uniform), the u/v face columns.
struct ScalCol {
    int i, j, rA, yo[7];
    unsigned cx[7], cu1;
};

__device__ __forceinline__ bool scal_col(const Dom& d, ScalCol& s)
{
    const RowSeg rs = row_seg(d.nx, d.ny);
    if (!rs.live) return false;
    const int nx = d.nx, ny = d.ny;
    s.i = rs.i;
    s.j = rs.j;
    s.rA = rs.j * nx;
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        s.yo[o] = pmod(rs.j + o - 3, ny) * nx;
        s.cx[o] = RB * pmod(rs.i + o - 3, nx);
    }
    s.cu1 = RB * (rs.i + 1);
    return true;
}

// Unlimited coupled stage update of up to NS species sharing the
// substep-averaged fluxes ru_m, rv_m, ww_m (the final stage clamps at
// zero when clamp is set):
//   q = (C(mu_t) q_t + dt div(q_cur; ru_m, rv_m, ww_m)) / C(mu_new).
// Same float32 expressions, in the same order, as the reference's flux
// divergence and update, so the result is bit-identical.
template <int NS>
__global__ __launch_bounds__(256) void k_scalar_cols(Dom d, QGroup g,
                                                     const real* RS mup0,
                                                     real dt, real rdx,
                                                     real rdy, int clamp)
{
    ScalCol q;
    if (!scal_col(d, q)) return;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    const int pu = nx + 1, stu = ny * pu, stv = (ny + 1) * nx;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* rdnw = kview(d.rdnw);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(q.cx[o]);
        colb_opaque(q.cu1);
    };
    const unsigned cC = q.cx[3];
    const int c = q.rA + q.i;
    const real mu0 = d.mub2d[c] + mup0[c];
    const real mu = d.mub2d[c] + (d.mup[c] + d.mu_pp[c]);
    real qz[NS][5], fz0[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        qz[s][0] = qz[s][1] = R(0.0);
        qz[s][2] = gld(g.cur[s], q.rA, cC);
        qz[s][3] = gld(g.cur[s], st + q.rA, cC);
        qz[s][4] = gld(g.cur[s], 2 * st + q.rA, cC);
        fz0[s] = R(0.0);                    // no flux through the surface
    }
    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const int lm = k * st, lu = k * stu + q.j * pu, lv = k * stv;
        const real ru0 = gld(d.ru_m, lu, cC), ru1 = gld(d.ru_m, lu, q.cu1);
        const real rv0 = gld(d.rv_m, lv + q.rA, cC);
        const real rv1 = gld(d.rv_m, lv + q.rA + nx, cC);
        const real ww1 = gld(d.ww_m, lm + st + q.rA, cC);
        const real chm0 = c1h[k] * mu0 + c2h[k];
        const real chm = c1h[k] * mu + c2h[k];
        const int kn = k + 3 < nz ? k + 3 : nz - 1;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            real x[7], y[7];
#pragma unroll
            for (int o = 0; o < 7; ++o) {
                x[o] = (o == 3) ? qz[s][2] : gld(g.cur[s], lm + q.rA, q.cx[o]);
                y[o] = (o == 3) ? qz[s][2] : gld(g.cur[s], lm + q.yo[o], cC);
            }
            const real fx0 = flux5(x[0], x[1], x[2], x[3], x[4], x[5], ru0);
            const real fx1 = flux5(x[1], x[2], x[3], x[4], x[5], x[6], ru1);
            const real fy0 = flux5(y[0], y[1], y[2], y[3], y[4], y[5], rv0);
            const real fy1 = flux5(y[1], y[2], y[3], y[4], y[5], y[6], rv1);
            const real fz1 = zf(qz[s][1], qz[s][2], qz[s][3], qz[s][4], ww1,
                                k + 1, nz, fnm, fnp);
            real tend = R(0.0) + div3(fx1 - fx0, fy1 - fy0,
                                      fz1 - fz0[s], rdx, rdy, rdnw[k]);
            if (g.fixed[s])              // held mixing (moist.py:384-386)
                tend = tend + gld(g.fixed[s], lm + q.rA, cC);
            fz0[s] = fz1;
            real v = chm0 * gld(g.q0[s], lm + q.rA, cC);
            const real t = dt * tend;
            v = v + t;
            v = v / chm;
            if (clamp && !(v > R(0.0)) && v == v) v = R(0.0);
            gst(g.dst[s], lm + q.rA, cC, v);
            const real qn = gld(g.cur[s], kn * st + q.rA, cC);
            qz[s][0] = qz[s][1];
            qz[s][1] = qz[s][2];
            qz[s][2] = qz[s][3];
            qz[s][3] = qz[s][4];
            qz[s][4] = (k + 3 < nz) ? qn : R(0.0);
        }
    }
}

// The positive-definite limiter's face fluxes of one cell (reference
// pd_advection.cu pd_fluxes, periodic, no msf): at each face the
// CFL-clamped upwind flux fl of the time-t scalar and the correction
// fc = (unlimited flux of the stage scalar) - fl.  x: faces i (l), i+1
// (r); y: rows j (l), j+1 (r).  The vertical faces are carried up the
// column by the caller.
struct PdFaces {
    real xl_l, xl_c, xr_l, xr_c, yl_l, yl_c, yr_l, yr_c;
};

struct PdCol {
    real mW, mA, mE, mS, mN;     // post-acoustic stage mass of the 5 columns
};

__device__ __forceinline__ PdFaces pd_faces(const Dom& d, const ScalCol& q,
                                            const PdCol& m,
                                            const real* RS cur,
                                            const real* RS q0, int k,
                                            real dx, real dy, real dt,
                                            real c1k, real c2k, real q0A,
                                            real qcA)
{
    const int ny = d.ny, nx = d.nx, st = ny * nx;
    const int pu = nx + 1, stu = ny * pu, stv = (ny + 1) * nx;
    const int lm = k * st, lu = k * stu + q.j * pu, lv = k * stv;
    const unsigned cC = q.cx[3];
    PdFaces f;
    real x[7], y[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) {
        x[o] = (o == 3) ? qcA : gld(cur, lm + q.rA, q.cx[o]);
        y[o] = (o == 3) ? qcA : gld(cur, lm + q.yo[o], cC);
    }
    const real q0W = gld(q0, lm + q.rA, q.cx[2]);
    const real q0E = gld(q0, lm + q.rA, q.cx[4]);
    const real q0S = gld(q0, lm + q.yo[2], cC);
    const real q0N = gld(q0, lm + q.yo[4], cC);
    {   // x face i: cells i-1 (upwind left) and i
        const real vel = gld(d.ru_m, lu, cC);
        const real muf = (c1k * R(0.5)) * (m.mA + m.mW) + c2k;
        const real cr = ((vel * dt) / dx) / muf;
        f.xl_l = (muf * (dx / dt)) * upwind_flux(q0W, q0A, cr);
        f.xl_c = flux5(x[0], x[1], x[2], x[3], x[4], x[5], vel) - f.xl_l;
    }
    {   // x face i+1: cells i and i+1
        const real vel = gld(d.ru_m, lu, q.cu1);
        const real muf = (c1k * R(0.5)) * (m.mE + m.mA) + c2k;
        const real cr = ((vel * dt) / dx) / muf;
        f.xr_l = (muf * (dx / dt)) * upwind_flux(q0A, q0E, cr);
        f.xr_c = flux5(x[1], x[2], x[3], x[4], x[5], x[6], vel) - f.xr_l;
    }
    {   // y face j
        const real vel = gld(d.rv_m, lv + q.rA, cC);
        const real muf = (c1k * R(0.5)) * (m.mA + m.mS) + c2k;
        const real cr = ((vel * dt) / dy) / muf;
        f.yl_l = (muf * (dy / dt)) * upwind_flux(q0S, q0A, cr);
        f.yl_c = flux5(y[0], y[1], y[2], y[3], y[4], y[5], vel) - f.yl_l;
    }
    {   // y face j+1
        const real vel = gld(d.rv_m, lv + q.rA + nx, cC);
        const real muf = (c1k * R(0.5)) * (m.mN + m.mA) + c2k;
        const real cr = ((vel * dt) / dy) / muf;
        f.yr_l = (muf * (dy / dt)) * upwind_flux(q0A, q0N, cr);
        f.yr_c = flux5(y[1], y[2], y[3], y[4], y[5], y[6], vel) - f.yr_l;
    }
    return f;
}

// Vertical face kf (1 <= kf <= nz-1) of the limiter: (fl, fc) from the
// time-t scalar below/above and the stage scalar's window qm2..qp1 around
// the face; zero at the surface and the lid.
__device__ __forceinline__ void pd_zface(const Dom& d, kreal* rdnw,
                                         kreal* c1h, kreal* c2h, kreal* fnm,
                                         kreal* fnp, int kf, real vel,
                                         real mA, real q0b, real q0t,
                                         real qm2, real qm1, real q0v,
                                         real qp1, real dt, real& fl,
                                         real& fc)
{
    const int nz = d.nz;
    const int kc = kf < nz ? kf : nz - 1, kb = kf > 0 ? kf - 1 : 0;
    const real dz = R(2.0) / (rdnw[kc] + rdnw[kb]);
    const real muf = c1h[kc] * mA + c2h[kc];
    const real cr = (vel * dt) / (dz * muf);
    const real l = (muf * (dz / dt)) * upwind_flux(q0b, q0t, cr);
    const real h = zf(qm2, qm1, q0v, qp1, vel, kf, nz, fnm, fnp);
    const bool in = kf > 0 && kf < nz;
    fl = in ? l : R(0.0);
    fc = in ? h - l : R(0.0);
}

__device__ __forceinline__ PdCol pd_col(const Dom& d, const ScalCol& q)
{
    auto mut = [&](int row, unsigned cb) {
        return gld(d.mub2d, row, cb)
               + (gld(d.mup, row, cb) + gld(d.mu_pp, row, cb));
    };
    PdCol m;
    m.mA = mut(q.rA, q.cx[3]);
    m.mW = mut(q.rA, q.cx[2]);
    m.mE = mut(q.rA, q.cx[4]);
    m.mS = mut(q.yo[2], q.cx[3]);
    m.mN = mut(q.yo[4], q.cx[3]);
    return m;
}

// Pass 1 of the final-stage limiter (reference pd_advection.cu pd_scale,
// periodic no-msf branch): the renormalisation factor of every cell from
// its face fluxes, recomputed here instead of stored.
__global__ __launch_bounds__(256) void k_pd_scale_cols(
    Dom d, const real* RS cur, const real* RS q0, const real* RS mup0,
    real* RS sc, real dx, real dy, real dxi, real dyi, real dt)
{
    ScalCol q;
    if (!scal_col(d, q)) return;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* rdnw = kview(d.rdnw);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(q.cx[o]);
        colb_opaque(q.cu1);
    };
    const unsigned cC = q.cx[3];
    const int c = q.rA + q.i;
    const PdCol m = pd_col(d, q);
    const real mu0 = d.mub2d[c] + mup0[c];
    // stage scalar window q(k-2..k+2), time-t scalar at k, k+1
    real w[5] = {R(0.0), R(0.0), gld(cur, q.rA, cC), gld(cur, st + q.rA, cC),
                 gld(cur, 2 * st + q.rA, cC)};
    real t0 = gld(q0, q.rA, cC), t1 = gld(q0, st + q.rA, cC);
    real zb_l = R(0.0), zb_c = R(0.0);      // surface face
    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const real c1k = c1h[k], c2k = c2h[k];
        const PdFaces f = pd_faces(d, q, m, cur, q0, k, dx, dy, dt, c1k, c2k,
                                   t0, w[2]);
        real zt_l, zt_c;
        const real vel = gld(d.ww_m, (k + 1) * st + q.rA, cC);
        pd_zface(d, rdnw, c1h, c2h, fnm, fnp, k + 1, vel, m.mA, t0, t1, w[1],
                 w[2], w[3], w[4], dt, zt_l, zt_c);
        const real chm0 = c1k * mu0 + c2k;
        const real dl = (dxi * (f.xr_l - f.xl_l) + dyi * (f.yr_l - f.yl_l))
                        + rdnw[k] * (zt_l - zb_l);
        const real ph_low = chm0 * t0 - dt * dl;
        const real fo = dt * ((dxi * (fmax(R(0.0), f.xr_c) - fmin(R(0.0), f.xl_c))
                               + dyi * (fmax(R(0.0), f.yr_c) - fmin(R(0.0), f.yl_c)))
                              + rdnw[k] * (fmin(R(0.0), zt_c) - fmax(R(0.0), zb_c)));
        real r = R(1.0);
        if (fo > ph_low) r = fmin(R(1.0), fmax(R(0.0), ph_low / (fo + R(1e-20))));
        gst(sc, k * st + q.rA, cC, r);
        zb_l = zt_l;
        zb_c = zt_c;
        const int kn = k + 3 < nz ? k + 3 : nz - 1;
        const real wn = gld(cur, kn * st + q.rA, cC);
        w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4];
        w[4] = (k + 3 < nz) ? wn : R(0.0);
        const int kt = k + 2 < nz ? k + 2 : nz - 1;
        const real tn = gld(q0, kt * st + q.rA, cC);
        t0 = t1;
        t1 = (k + 2 < nz) ? tn : R(0.0);
    }
}

// Pass 2 (reference pd_advection.cu pd_renorm_apply + moist.py:385-398,
// final stage, clamped): every correction flux takes its donor cell's
// factor; the coupled update goes to out (not q0's array, whose
// neighbours the face fluxes still read).
__global__ __launch_bounds__(256) void k_pd_apply_cols(
    Dom d, const real* RS cur, const real* RS q0, const real* RS mup0,
    const real* RS sc, real* RS out, real dx, real dy, real dxi, real dyi,
    real dt)
{
    ScalCol q;
    if (!scal_col(d, q)) return;
    const int nz = d.nz, ny = d.ny, nx = d.nx, st = ny * nx;
    kreal* c1h = kview(d.c1h);
    kreal* c2h = kview(d.c2h);
    kreal* rdnw = kview(d.rdnw);
    kreal* fnm = kview(d.fnm);
    kreal* fnp = kview(d.fnp);
    auto opaque_all = [&]() {
#pragma unroll
        for (int o = 0; o < 7; ++o) colb_opaque(q.cx[o]);
        colb_opaque(q.cu1);
    };
    const unsigned cC = q.cx[3];
    const int c = q.rA + q.i;
    const PdCol m = pd_col(d, q);
    const real mu0 = d.mub2d[c] + mup0[c];
    const real mu = d.mub2d[c] + (d.mup[c] + d.mu_pp[c]);
    real w[5] = {R(0.0), R(0.0), gld(cur, q.rA, cC), gld(cur, st + q.rA, cC),
                 gld(cur, 2 * st + q.rA, cC)};
    real t0 = gld(q0, q.rA, cC), t1 = gld(q0, st + q.rA, cC);
    real zb_l = R(0.0), zb_c = R(0.0);
    real s_b = R(1.0), s0 = gld(sc, q.rA, cC);   // factors at k-1, k
    for (int k = 0; k < nz; ++k) {
        opaque_all();
        const int lm = k * st;
        const real c1k = c1h[k], c2k = c2h[k];
        const PdFaces f = pd_faces(d, q, m, cur, q0, k, dx, dy, dt, c1k, c2k,
                                   t0, w[2]);
        real zt_l, zt_c;
        const real vel = gld(d.ww_m, (k + 1) * st + q.rA, cC);
        pd_zface(d, rdnw, c1h, c2h, fnm, fnp, k + 1, vel, m.mA, t0, t1, w[1],
                 w[2], w[3], w[4], dt, zt_l, zt_c);
        const int ku = k + 1 < nz ? k + 1 : nz - 1;
        const real s_t = gld(sc, ku * st + q.rA, cC);
        const real sW = gld(sc, lm + q.rA, q.cx[2]);
        const real sE = gld(sc, lm + q.rA, q.cx[4]);
        const real sS = gld(sc, lm + q.yo[2], cC);
        const real sN = gld(sc, lm + q.yo[4], cC);
        const real sx_l = (f.xl_c > R(0.0)) ? sW : ((f.xl_c < R(0.0)) ? s0 : R(1.0));
        const real sx_r = (f.xr_c > R(0.0)) ? s0 : ((f.xr_c < R(0.0)) ? sE : R(1.0));
        const real sy_l = (f.yl_c > R(0.0)) ? sS : ((f.yl_c < R(0.0)) ? s0 : R(1.0));
        const real sy_r = (f.yr_c > R(0.0)) ? s0 : ((f.yr_c < R(0.0)) ? sN : R(1.0));
        const real sz_b = (zb_c > R(0.0)) ? s0 : ((zb_c < R(0.0)) ? s_b : R(1.0));
        const real sz_t = (zt_c < R(0.0)) ? s0 : ((zt_c > R(0.0)) ? s_t : R(1.0));
        const real gx = (sx_r * f.xr_c + f.xr_l) - (sx_l * f.xl_c + f.xl_l);
        const real gy = (sy_r * f.yr_c + f.yr_l) - (sy_l * f.yl_c + f.yl_l);
        const real gz = (sz_t * zt_c + zt_l) - (sz_b * zb_c + zb_l);
        const real tend = R(0.0) + (((-dxi) * gx - dyi * gy) - rdnw[k] * gz);
        const real chm0 = c1k * mu0 + c2k;
        const real chm = c1k * mu + c2k;
        real v = chm0 * t0;
        const real tt = dt * tend;
        v = v + tt;
        v = v / chm;
        if (!(v > R(0.0)) && v == v) v = R(0.0);
        gst(out, lm + q.rA, cC, v);
        zb_l = zt_l;
        zb_c = zt_c;
        s_b = s0;
        s0 = s_t;
        const int kn = k + 3 < nz ? k + 3 : nz - 1;
        const real wn = gld(cur, kn * st + q.rA, cC);
        w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4];
        w[4] = (k + 3 < nz) ? wn : R(0.0);
        const int kt = k + 2 < nz ? k + 2 : nz - 1;
        const real tn = gld(q0, kt * st + q.rA, cC);
        t0 = t1;
        t1 = (k + 2 < nz) ? tn : R(0.0);
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

// Whether a step option reaches the stage and substep kernels (moist_cq,
// h_sca_adv_order = 5, hypsometric_opt = 2, the retained heating, loading
// extra species): the plain step launches the variants compiled without
// them (XO = false), which keeps their registers and instruction count
// those of the step without options.
static inline bool xo_active(const Dom& d, const Sc& s)
{
    return s.prec || s.cq || s.h5 || d.hypsometric_opt == 2
           || d.h_diabatic != nullptr || d.n_load > 0;
}

// Calls f(std::true_type) when the options are live, else f(false_type).
template <class F>
static void with_xo(bool xo, F&& f)
{
    if (xo) f(std::true_type{});
    else f(std::false_type{});
}

template <bool PREC, bool XO>
static void launch_omega(const Dom& d, const Sc& s, int do_eos, hipStream_t st)
{
    const bool b3 = d.base3d != 0;
    if (do_eos && b3)
        hipLaunchKernelGGL((k_omega<One, true, true, PREC, XO>), cols(d), dim3(256), 0, st, d, s);
    else if (do_eos)
        hipLaunchKernelGGL((k_omega<One, true, false, PREC, XO>), cols(d), dim3(256), 0, st, d, s);
    else
        hipLaunchKernelGGL((k_omega<One, false, false, PREC, XO>), cols(d), dim3(256), 0, st, d, s);
}

hipError_t run_omega(const Dom& d, const Sc& s, int do_eos, hipStream_t st)
{
    if (s.prec) launch_omega<true, true>(d, s, do_eos, st);
    else if (xo_active(d, s)) launch_omega<false, true>(d, s, do_eos, st);
    else launch_omega<false, false>(d, s, do_eos, st);
    return hipGetLastError();
}

// Calls f(std::true_type) for the precision form, else f(std::false_type);
// the FP64 build has no precision form, so it instantiates only the
// latter.
template <class F>
static void with_prec(bool prec, F&& f)
{
    if constexpr (!ARW_F64) {
        if (prec) {
            f(std::true_type{});
            return;
        }
    }
    (void)prec;
    f(std::false_type{});
}

// Unused LDS reserved by every k_tend_faces launch: 40 KB of the CU's
// 160 KB admits at most 4 workgroups (4 waves per SIMD).  At 5 waves per
// SIMD the 7-row stencils' L2 reuse drops (13.4 instead of 11.3 fetched
// words per point at 1000x1000x50) and the kernel runs slower.
constexpr size_t kTendFacesLds = 40 * 1024;

template <class SRC, bool PREC, bool XO = true>
static void launch_tend_faces(const typename SRC::Arg& a, dim3 g, const Sc& s,
                              bool b3, hipStream_t st)
{
    const dim3 b(64, 4);
    const size_t l = kTendFacesLds;
    if (b3 && s.diff)
        hipLaunchKernelGGL((k_tend_faces<SRC, true, true, PREC, false, XO>), g, b, l, st, a, s);
    else if (b3)
        hipLaunchKernelGGL((k_tend_faces<SRC, true, false, PREC, false, XO>), g, b, l, st, a, s);
    else if (s.diff)
        hipLaunchKernelGGL((k_tend_faces<SRC, false, true, PREC, false, XO>), g, b, l, st, a, s);
    else
        hipLaunchKernelGGL((k_tend_faces<SRC, false, false, PREC, false, XO>), g, b, l, st, a, s);
}

template <class SRC, bool MOIST, bool PREC, bool XO = true>
static void launch_tend_cols(const typename SRC::Arg& a, dim3 g, const Sc& s,
                             bool b3, hipStream_t st)
{
    const dim3 b(64, 4);
    if (b3 && s.diff)
        hipLaunchKernelGGL((k_tend_cols<SRC, true, true, MOIST, PREC, false, XO>), g, b, 0, st, a, s);
    else if (b3)
        hipLaunchKernelGGL((k_tend_cols<SRC, true, false, MOIST, PREC, false, XO>), g, b, 0, st, a, s);
    else if (s.diff)
        hipLaunchKernelGGL((k_tend_cols<SRC, false, true, MOIST, PREC, false, XO>), g, b, 0, st, a, s);
    else
        hipLaunchKernelGGL((k_tend_cols<SRC, false, false, MOIST, PREC, false, XO>), g, b, 0, st, a, s);
}

// g: the whole launch; nmain: its leading column workgroups (one domain:
// the duplicate-face workgroups follow them; batched: unused)
template <class SRC, bool PREC, bool XO = true>
static void launch_finish(const typename SRC::Arg& a, dim3 g,
                          const FinishDst& o, bool b3, bool eos,
                          hipStream_t st, real hdt = R(0.0), int nmain = 0)
{
    if (b3 && eos)
        hipLaunchKernelGGL((k_finish<SRC, true, true, PREC, XO>), g, dim3(256), 0, st, a, o, hdt, nmain);
    else if (b3)
        hipLaunchKernelGGL((k_finish<SRC, true, false, PREC, XO>), g, dim3(256), 0, st, a, o, hdt, nmain);
    else if (eos)
        hipLaunchKernelGGL((k_finish<SRC, false, true, PREC, XO>), g, dim3(256), 0, st, a, o, hdt, nmain);
    else
        hipLaunchKernelGGL((k_finish<SRC, false, false, PREC, XO>), g, dim3(256), 0, st, a, o, hdt, nmain);
}

// The held-mixing variants: one domain, plain form (arw_step's held
// tendencies refuse the precision form and the ensemble).
template <bool B3, bool DIFF>
static void launch_tend_faces_held(const Dom& d, dim3 g, const Sc& s,
                                   hipStream_t st)
{
    hipLaunchKernelGGL((k_tend_faces<One, B3, DIFF, false, true>), g,
                       dim3(64, 4), kTendFacesLds, st, d, s);
}

template <bool B3, bool DIFF>
static void launch_tend_cols_held(const Dom& d, dim3 g, const Sc& s,
                                  hipStream_t st)
{
    if (d.qv)
        hipLaunchKernelGGL((k_tend_cols<One, B3, DIFF, true, false, true>), g,
                           dim3(64, 4), 0, st, d, s);
    else
        hipLaunchKernelGGL((k_tend_cols<One, B3, DIFF, false, false, true>), g,
                           dim3(64, 4), 0, st, d, s);
}

hipError_t run_tend_faces(const Dom& d, const Sc& s, hipStream_t st)
{
    const dim3 g = rowseg_grid(d.nx, d.ny);
    const bool b3 = d.base3d != 0;
    if (s.held && !s.prec) {
        if (b3 && s.diff) launch_tend_faces_held<true, true>(d, g, s, st);
        else if (b3) launch_tend_faces_held<true, false>(d, g, s, st);
        else if (s.diff) launch_tend_faces_held<false, true>(d, g, s, st);
        else launch_tend_faces_held<false, false>(d, g, s, st);
    } else if (xo_active(d, s)) {
        with_prec(s.prec, [&](auto p) {
            launch_tend_faces<One, decltype(p)::value>(d, g, s, b3, st);
        });
    } else {
        launch_tend_faces<One, false, false>(d, g, s, b3, st);
    }
    const int n = d.nz * d.ny + d.nz * d.nx;
    hipLaunchKernelGGL(k_tend_faces_edge<One>, dim3((n + 255) / 256), dim3(256), 0,
                       st, d);
    return hipGetLastError();
}

hipError_t run_tend_cols(const Dom& d, const Sc& s, hipStream_t st)
{
    const dim3 g = rowseg_grid(d.nx, d.ny);
    const bool b3 = d.base3d != 0;
    if (s.held && !s.prec) {
        if (b3 && s.diff) launch_tend_cols_held<true, true>(d, g, s, st);
        else if (b3) launch_tend_cols_held<true, false>(d, g, s, st);
        else if (s.diff) launch_tend_cols_held<false, true>(d, g, s, st);
        else launch_tend_cols_held<false, false>(d, g, s, st);
        return hipGetLastError();
    }
    if (!xo_active(d, s)) {
        if (d.qv) launch_tend_cols<One, true, false, false>(d, g, s, b3, st);
        else launch_tend_cols<One, false, false, false>(d, g, s, b3, st);
        return hipGetLastError();
    }
    with_prec(s.prec, [&](auto p) {
        constexpr bool P = decltype(p)::value;
        if (d.qv) launch_tend_cols<One, true, P>(d, g, s, b3, st);
        else launch_tend_cols<One, false, P>(d, g, s, b3, st);
    });
    return hipGetLastError();
}

template <bool B3, bool USEP, bool XO>
static void launch_uv(const Dom& d, const Sc& s, hipStream_t st)
{
    const int ne = (int)xcd_grid((unsigned)((d.ny + d.nx + 255) / 256));
    const dim3 g(rowseg_grid(d.nx, d.ny).x + ne);
    if (s.sum)
        hipLaunchKernelGGL((k_uv<One, B3, USEP, true, XO>), g, dim3(64, 4), 0,
                           st, d, s, ne);
    else
        hipLaunchKernelGGL((k_uv<One, B3, USEP, false, XO>), g, dim3(64, 4), 0,
                           st, d, s, ne);
}

hipError_t run_uv(const Dom& d, const Sc& s, hipStream_t st)
{
    const bool b3 = d.base3d != 0, usep = s.use_pold != 0;
    with_xo(xo_active(d, s), [&](auto x) {
        constexpr bool X = decltype(x)::value;
        if (b3 && usep) launch_uv<true, true, X>(d, s, st);
        else if (b3) launch_uv<true, false, X>(d, s, st);
        else if (usep) launch_uv<false, true, X>(d, s, st);
        else launch_uv<false, false, X>(d, s, st);
    });
    return hipGetLastError();
}

// Levels per column whose Thomas factors and eliminated w'' k_column keeps
// in LDS: 128 bytes per column (FP32), which leaves its VGPR-bound
// occupancy unchanged; the precision form parks its FP64 pair in HBM.
static inline Sc column_sc(const Sc& s)
{
    Sc c = s;
    c.park = s.prec ? 0 : (int)(128 / (2 * sizeof(real)));
    return c;
}

static inline size_t park_bytes(const Sc& c, int cols_per_block)
{
    return (size_t)2 * c.park * cols_per_block * sizeof(real);
}

template <bool B3, bool DAMP, bool AL, bool WW>
static void launch_column(const Dom& d, const Sc& s0, real* pout,
                          hipStream_t st)
{
    const Sc s = column_sc(s0);
    const size_t lds = park_bytes(s, 256);
    if (s.prec)
        hipLaunchKernelGGL((k_column<One, B3, DAMP, AL, WW, true>), cols(d),
                           dim3(256), lds, st, d, s, pout);
    else if (xo_active(d, s))
        hipLaunchKernelGGL((k_column<One, B3, DAMP, AL, WW, false>), cols(d),
                           dim3(256), lds, st, d, s, pout);
    else
        hipLaunchKernelGGL((k_column<One, B3, DAMP, AL, WW, false, false>),
                           cols(d), dim3(256), lds, st, d, s, pout);
}

template <bool B3, bool DAMP>
static void launch_column_bd(const Dom& d, const Sc& s, real* pout, int al,
                             int ww, hipStream_t st)
{
    if (al && ww) launch_column<B3, DAMP, true, true>(d, s, pout, st);
    else if (al) launch_column<B3, DAMP, true, false>(d, s, pout, st);
    else if (ww) launch_column<B3, DAMP, false, true>(d, s, pout, st);
    else launch_column<B3, DAMP, false, false>(d, s, pout, st);
}

hipError_t run_column(const Dom& d, const Sc& s, real* pout, int write_al,
                      int write_ww, hipStream_t st)
{
    const bool b3 = d.base3d != 0, damp = s.damp != 0;
    if (b3 && damp) launch_column_bd<true, true>(d, s, pout, write_al, write_ww, st);
    else if (b3) launch_column_bd<true, false>(d, s, pout, write_al, write_ww, st);
    else if (damp) launch_column_bd<false, true>(d, s, pout, write_al, write_ww, st);
    else launch_column_bd<false, false>(d, s, pout, write_al, write_ww, st);
    return hipGetLastError();
}

hipError_t run_finish(const Dom& d, const FinishDst& o, bool prec,
                      hipStream_t st, double hdiab_dt)
{
    // a moist state's EOS waits for the scalar update of the stage
    const bool eos = d.qv == nullptr;
    const real hdt = d.h_diabatic ? (real)hdiab_dt : R(0.0);
    const bool xo = prec || d.hypsometric_opt == 2 || d.h_diabatic != nullptr;
    // the column workgroups, then those of the duplicate faces
    const int nmain = (int)cols(d).x;
    const dim3 g(nmain + (d.ny + d.nx + 255) / 256);
    if (xo) {
        with_prec(prec, [&](auto p) {
            launch_finish<One, decltype(p)::value>(d, g, o, d.base3d != 0,
                                                   eos, st, hdt, nmain);
        });
    } else {
        launch_finish<One, false, false>(d, g, o, d.base3d != 0, eos, st,
                                         hdt, nmain);
    }
    return hipGetLastError();
}

hipError_t run_scalar_cols(const Dom& d, const QGroup& g, const real* mup0,
                           real dt, real rdx, real rdy, int clamp,
                           hipStream_t st)
{
    const dim3 grid = rowseg_grid(d.nx, d.ny), blk(64, 4);
    switch (g.n) {
    case 1:
        hipLaunchKernelGGL(k_scalar_cols<1>, grid, blk, 0, st, d, g, mup0, dt,
                           rdx, rdy, clamp);
        break;
    case 2:
        hipLaunchKernelGGL(k_scalar_cols<2>, grid, blk, 0, st, d, g, mup0, dt,
                           rdx, rdy, clamp);
        break;
    case 3:
        hipLaunchKernelGGL(k_scalar_cols<3>, grid, blk, 0, st, d, g, mup0, dt,
                           rdx, rdy, clamp);
        break;
    case 4:
        hipLaunchKernelGGL(k_scalar_cols<4>, grid, blk, 0, st, d, g, mup0, dt,
                           rdx, rdy, clamp);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t run_pd_cols(const Dom& d, const real* cur, const real* q0,
                       real* out, const real* mup0, real dx, real dy,
                       real dxi, real dyi, real dt, hipStream_t st)
{
    const dim3 grid = rowseg_grid(d.nx, d.ny), blk(64, 4);
    hipLaunchKernelGGL(k_pd_scale_cols, grid, blk, 0, st, d, cur, q0, mup0,
                       d.pd_sc, dx, dy, dxi, dyi, dt);
    hipLaunchKernelGGL(k_pd_apply_cols, grid, blk, 0, st, d, cur, q0, mup0,
                       d.pd_sc, out, dx, dy, dxi, dyi, dt);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched launchers (Many)
// ---------------------------------------------------------------------------

int batch_blocks(const Dom& d, int kind, int col_tile)
{
    const int rows = ((d.nx + 63) / 64) * ((d.ny + 3) / 4);
    switch (kind) {
    case BK_COLS: return (d.ny * d.nx + 255) / 256;
    case BK_COLT: return (d.ny * d.nx + col_tile - 1) / col_tile;
    case BK_ROWS: return rows;
    case BK_UV: return (d.ny + d.nx + 255) / 256 + rows;
    case BK_TFE: return (d.nz * d.ny + d.nz * d.nx + 255) / 256;
    default: return (d.ny + d.nx + 255) / 256;
    }
}

static inline dim3 bgrid(const Batch& b)
{
    return dim3(xcd_grid((unsigned)b.total));
}

template <bool PREC>
static void launch_omega_b(const Batch& b, const Sc& s, int do_eos, bool b3,
                           hipStream_t st)
{
    if (do_eos && b3)
        hipLaunchKernelGGL((k_omega<Many, true, true, PREC>), bgrid(b), dim3(256), 0, st, b, s);
    else if (do_eos)
        hipLaunchKernelGGL((k_omega<Many, true, false, PREC>), bgrid(b), dim3(256), 0, st, b, s);
    else
        hipLaunchKernelGGL((k_omega<Many, false, false, PREC>), bgrid(b), dim3(256), 0, st, b, s);
}

hipError_t run_omega_b(const BatchSet& b, const Sc& s, int do_eos, bool b3,
                       hipStream_t st)
{
    if (s.prec) launch_omega_b<true>(b.k[BK_COLS], s, do_eos, b3, st);
    else launch_omega_b<false>(b.k[BK_COLS], s, do_eos, b3, st);
    return hipGetLastError();
}

hipError_t run_tend_faces_b(const BatchSet& b, const Sc& s, bool b3,
                            hipStream_t st)
{
    const Batch& r = b.k[BK_ROWS];
    with_prec(s.prec, [&](auto p) {
        launch_tend_faces<Many, decltype(p)::value>(r, bgrid(r), s, b3, st);
    });
    const Batch& e = b.k[BK_TFE];
    hipLaunchKernelGGL(k_tend_faces_edge<Many>, bgrid(e), dim3(256), 0, st, e);
    return hipGetLastError();
}

hipError_t run_tend_cols_b(const BatchSet& b, const Sc& s, bool b3,
                           hipStream_t st)
{
    const Batch& r = b.k[BK_ROWS];
    with_prec(s.prec, [&](auto p) {
        launch_tend_cols<Many, false, decltype(p)::value>(r, bgrid(r), s, b3, st);
    });
    return hipGetLastError();
}

hipError_t run_uv_b(const BatchSet& b, const Sc& s, bool b3, hipStream_t st)
{
    const Batch& r = b.k[BK_UV];
    const dim3 g = bgrid(r), blk(64, 4);
    const bool usep = s.use_pold != 0;
    if (b3 && usep)
        hipLaunchKernelGGL((k_uv<Many, true, true>), g, blk, 0, st, r, s, 0);
    else if (b3)
        hipLaunchKernelGGL((k_uv<Many, true, false>), g, blk, 0, st, r, s, 0);
    else if (usep)
        hipLaunchKernelGGL((k_uv<Many, false, true>), g, blk, 0, st, r, s, 0);
    else
        hipLaunchKernelGGL((k_uv<Many, false, false>), g, blk, 0, st, r, s, 0);
    return hipGetLastError();
}

template <bool B3, bool DAMP, bool AL>
static void launch_column_b(const Batch& r, const Sc& s0, hipStream_t st)
{
    const Sc s = column_sc(s0);
    const size_t lds = park_bytes(s, r.tile);
    if (s.prec)
        hipLaunchKernelGGL((k_column<Many, B3, DAMP, AL, false, true>),
                           bgrid(r), dim3(r.tile), lds, st, r, s, nullptr);
    else
        hipLaunchKernelGGL((k_column<Many, B3, DAMP, AL, false, false>),
                           bgrid(r), dim3(r.tile), lds, st, r, s, nullptr);
}

template <bool B3, bool DAMP>
static void launch_column_b_al(const Batch& r, const Sc& s, int al,
                               hipStream_t st)
{
    if (al) launch_column_b<B3, DAMP, true>(r, s, st);
    else launch_column_b<B3, DAMP, false>(r, s, st);
}

hipError_t run_column_b(const BatchSet& b, const Sc& s, bool b3, int write_al,
                        hipStream_t st)
{
    const Batch& r = b.k[BK_COLT];
    const bool damp = s.damp != 0;
    if (b3 && damp) launch_column_b_al<true, true>(r, s, write_al, st);
    else if (b3) launch_column_b_al<true, false>(r, s, write_al, st);
    else if (damp) launch_column_b_al<false, true>(r, s, write_al, st);
    else launch_column_b_al<false, false>(r, s, write_al, st);
    return hipGetLastError();
}

hipError_t run_finish_b(const BatchSet& b, bool b3, bool prec, hipStream_t st)
{
    const Batch& r = b.k[BK_COLS];
    const FinishDst none = {};
    with_prec(prec, [&](auto p) {
        launch_finish<Many, decltype(p)::value>(r, bgrid(r), none, b3, true, st);
    });
    const Batch& e = b.k[BK_FE];
    hipLaunchKernelGGL(k_finish_edge<Many>, bgrid(e), dim3(256), 0, st, e, none);
    return hipGetLastError();
}

}  // namespace ARW_NS
