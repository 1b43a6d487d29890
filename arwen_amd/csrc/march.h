// Launchers of the k-marching step kernels (march.hip).
#pragma once

#include "common.h"

namespace ARW_NS {

// New-state destinations of the small-step finish (stage buffers rotate
// instead of copying the time-t fields).
struct FinishDst {
    real *u, *v, *w, *thp, *php, *mup;
};

// One launch over every member of an ensemble (see Many below); doms (and
// fin for the finish) hold the members' views for this launch, owner and
// start the workgroup tables of its launch shape.  Device memory.
struct Batch {
    const Dom* doms;
    const FinishDst* fin;
    const int* owner;
    const int* start;
    int total;
    int tile;          // threads per workgroup of this launch shape
};

// Workgroups one member needs in each batched launch shape.  BK_COLT is
// the column solve's: col_tile columns per workgroup (64, 128 or 256,
// chosen per ensemble); the other column launches keep 256.
enum BatchKind { BK_COLS, BK_ROWS, BK_UV, BK_TFE, BK_FE, BK_COLT, BK_COUNT };
int batch_blocks(const Dom& d, int kind, int col_tile);

// The members' views for one launch role, with the per-shape tables.
struct BatchSet {
    Batch k[BK_COUNT];
};

// ---------------------------------------------------------------------------
// Where a kernel finds its domain.  One: the kernel argument, workgroup
// blockIdx.x.  Many: a member of an ensemble launch (Batch) -- logical
// workgroup g of the launch (XCD-grouped) belongs to member owner[g] and
// is that member's workgroup g - start[owner[g]].  The member views are
// read through the constant address space, like kernel arguments: scalar
// loads the compiler may repeat anywhere.
// ---------------------------------------------------------------------------
typedef const __attribute__((address_space(4))) int kint;
typedef const __attribute__((address_space(4))) Dom kdom;
typedef const __attribute__((address_space(4))) FinishDst kfin;

struct One {
    static constexpr bool many = false;
    typedef Dom Arg;
    static __device__ __forceinline__ bool locate(const Dom&, int& m, int& g)
    {
        m = 0;
        g = (int)blockIdx.x;
        return true;
    }
    static __device__ __forceinline__ const Dom& dom(const Dom& a, int)
    {
        return a;
    }
};

struct Many {
    static constexpr bool many = true;
    typedef Batch Arg;
    static __device__ __forceinline__ bool locate(const Batch& b, int& m,
                                                  int& g)
    {
        const int x = xcd_block();
        if (x >= b.total) return false;
        m = ((kint*)b.owner)[x];
        g = x - ((kint*)b.start)[m];
        return true;
    }
    static __device__ __forceinline__ const Dom& dom(const Batch& b, int m)
    {
        return *(const Dom*)((kdom*)b.doms + m);
    }
    static __device__ __forceinline__ const FinishDst& fin(const Batch& b,
                                                           int m)
    {
        return *(const FinishDst*)((kfin*)b.fin + m);
    }
};

template <class SRC>
__device__ __forceinline__ const FinishDst& fin_of(const Dom&, int,
                                                   const FinishDst& o)
{
    return o;
}

template <class SRC>
__device__ __forceinline__ const FinishDst& fin_of(const Batch& b, int m,
                                                   const FinishDst&)
{
    return Many::fin(b, m);
}

// Error reporting and argument checks shared by the library's sources
// (defined in dycore.hip).
int32_t lib_fail(int32_t code, const char* msg);

// WRF km_opt = 4 held tendencies of the time-t fields of d into out_u,
// out_v, out_w, out_t and of the nq species q_src into held_q (smag.hip);
// buf: 9 nz ny nx + 5 (nz+1) ny nx elements of scratch
hipError_t run_smag(const Dom& d, const arw_params* prm, real* buf,
                    real* out_u, real* out_v, real* out_w, real* out_t,
                    const real* const* q_src, real* held_q, int nq,
                    hipStream_t st);
int32_t lib_launched(hipError_t e, const char* what);
int32_t lib_check_domain(const Dom* d);
int32_t lib_check_adj_domain(const Dom* d);
int32_t lib_check_prec(const arw_params* p, const Dom* d);

hipError_t run_omega(const Dom& d, const Sc& s, int do_eos, hipStream_t st);
// The plain EOS of the state into d.p, d.al, d.alt (dycore.hip, k_eos).
hipError_t run_eos(const Dom& d, hipStream_t st);
// Coupled mass fluxes ru, rv of the state (dycore.hip, k_faces).
hipError_t run_faces(const Dom& d, real* ru, real* rv, hipStream_t st);
// The advection adjoints of advect_adj.hip (op 0 scalar, 1 u, 2 v, 3 w),
// adding into g_q, g_ru, g_rv, g_rw.
int32_t lib_flux_div_adjoint(int op, const real* q, const real* ru,
                             const real* rv, const real* rw, const real* lam,
                             real* g_q, real* g_ru, real* g_rv, real* g_rw,
                             const real* rdz, const real* fnm,
                             const real* fnp, real dx_inv, real dy_inv,
                             int nz, int ny, int nx, hipStream_t st);
// Up to four scalars of one stage update that share the advecting fluxes.
struct QGroup {
    const real* cur[4];
    const real* q0[4];
    real* dst[4];
    const real* fixed[4];   // held mixing tendencies, or null
    int n;
};

// Scalar transport of one RK stage as column marches: the unlimited
// coupled update of a group, and the final stage's positive-definite
// limiter of one scalar (factors in d.pd_sc, the update into out).
hipError_t run_scalar_cols(const Dom& d, const QGroup& g, const real* mup0,
                           real dt, real rdx, real rdy, int clamp,
                           hipStream_t st);
hipError_t run_pd_cols(const Dom& d, const real* cur, const real* q0,
                       real* out, const real* mup0, real dx, real dy,
                       real dxi, real dyi, real dt, hipStream_t st);

hipError_t run_tend_faces(const Dom& d, const Sc& s, hipStream_t st);
hipError_t run_tend_cols(const Dom& d, const Sc& s, hipStream_t st);
hipError_t run_uv(const Dom& d, const Sc& s, hipStream_t st);
hipError_t run_column(const Dom& d, const Sc& s, real* pout, int write_al,
                      int write_ww, hipStream_t st);
hipError_t run_finish(const Dom& d, const FinishDst& o, bool prec,
                      hipStream_t st, double hdiab_dt = 0.0);

// Batched forms (members share base3d and every flag in Sc).
hipError_t run_omega_b(const BatchSet& b, const Sc& s, int do_eos, bool b3,
                       hipStream_t st);
hipError_t run_tend_faces_b(const BatchSet& b, const Sc& s, bool b3,
                            hipStream_t st);
hipError_t run_tend_cols_b(const BatchSet& b, const Sc& s, bool b3,
                           hipStream_t st);
hipError_t run_uv_b(const BatchSet& b, const Sc& s, bool b3, hipStream_t st);
hipError_t run_column_b(const BatchSet& b, const Sc& s, bool b3, int write_al,
                        hipStream_t st);
hipError_t run_finish_b(const BatchSet& b, bool b3, bool prec, hipStream_t st);

}  // namespace ARW_NS
