// Launchers of the k-marching step kernels (march.hip).
#pragma once

#include "common.h"

namespace ARW_NS {

// New-state destinations of the small-step finish (stage buffers rotate
// instead of copying the time-t fields).
struct FinishDst {
    real *u, *v, *w, *thp, *php, *mup;
};

hipError_t run_omega(const Dom& d, const Sc& s, int do_eos, hipStream_t st);
hipError_t run_tend_faces(const Dom& d, const Sc& s, hipStream_t st);
hipError_t run_tend_cols(const Dom& d, const Sc& s, hipStream_t st);
hipError_t run_uv(const Dom& d, const Sc& s, hipStream_t st);
hipError_t run_column(const Dom& d, const Sc& s, real* pout, int write_al,
                      int write_ww, hipStream_t st);
hipError_t run_finish(const Dom& d, const FinishDst& o, hipStream_t st);

}  // namespace ARW_NS
