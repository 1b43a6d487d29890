// Full-state health gate (reference gpuwm/core/health.py:611-770,
// kernels/health.cu:224-297): one launch checks every registered field of
// the state against its bounds (finite; lower, optionally strict; upper),
// optionally on the field plus an auxiliary (theta' + the base theta, mu'
// + the base column mass), and leaves a two-word record: the OR of the
// failing fields' status bits and the lexicographically first failing
// (field number, flat index) packed as field << 48 | index.
//
// Launch shape: blockIdx.y is the field, blockIdx.x a chunk of it, so the
// largest field does not leave the rest of the device idle.  Each 64-lane
// wave reduces its first failing index with a cross-lane minimum (all ones
// when nothing failed), then one lane per wave folds it into the record
// with 64-bit atomics.  A tiny first kernel writes the record's identity
// (0, all ones), so the pair is safe to capture in a graph (no memset).
#include "common.h"

namespace ARW_NS {

namespace {

constexpr uint32_t kLower = 1u << 0, kUpper = 1u << 1, kStrictLower = 1u << 2,
                   kAddAux = 1u << 3, kAuxLevel = 1u << 4;

__global__ void k_health_init(unsigned long long* result)
{
    if (threadIdx.x == 0) {
        result[0] = 0ull;
        result[1] = ~0ull;
    }
}

__device__ __forceinline__ unsigned long long wave_min(unsigned long long x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(x, off, 64);
        x = o < x ? o : x;
    }
    return x;
}

__global__ __launch_bounds__(256) void k_health_gate(
    const arw_health_field* __restrict__ fields, int nfields,
    unsigned long long chunk, unsigned long long* __restrict__ result)
{
    const int fi = (int)blockIdx.y;
    if (fi >= nfields) return;
    const arw_health_field f = fields[fi];
    const unsigned long long size = f.size;
    const unsigned long long begin = (unsigned long long)blockIdx.x * chunk;
    if (begin >= size) return;
    const unsigned long long end = begin + chunk < size ? begin + chunk : size;
    // offsets within the chunk are 32-bit; the level of an element (for a
    // per-level auxiliary) is carried along instead of divided out
    const real* __restrict__ v = static_cast<const real*>(f.values) + begin;
    const real* __restrict__ aux = static_cast<const real*>(f.aux);
    const real lo = (real)f.lower, hi = (real)f.upper;
    const uint32_t flags = f.flags;
    const unsigned n = (unsigned)(end - begin), step = blockDim.x;
    const unsigned plane = f.plane ? f.plane : 1u;
    const bool level = (flags & kAuxLevel) != 0, add = (flags & kAddAux) != 0;
    unsigned lev = 0, rem = 0;
    if (level) {
        const unsigned long long ix0 = begin + threadIdx.x;
        lev = (unsigned)(ix0 / plane);
        rem = (unsigned)(ix0 - (unsigned long long)lev * plane);
    }
    const real* __restrict__ auxd = add && !level ? aux + begin : aux;
    unsigned first = ~0u;
    for (unsigned t = threadIdx.x; t < n; t += step) {
        real x = v[t];
        if (add) x = x + (level ? aux[lev] : auxd[t]);
        bool bad = !isfinite(x);
        if (flags & kLower) bad |= (flags & kStrictLower) ? x <= lo : x < lo;
        if (flags & kUpper) bad |= x > hi;
        if (bad && t < first) first = t;
        if (level) {
            rem += step;
            while (rem >= plane) {
                rem -= plane;
                ++lev;
            }
        }
    }
    const unsigned long long m =
        wave_min(first == ~0u ? ~0ull : begin + first);
    if ((threadIdx.x & 63) == 0 && m != ~0ull) {
        atomicOr(result, (unsigned long long)f.status_bit);
        atomicMin(result + 1, ((unsigned long long)fi << 48) | m);
    }
}

}  // namespace

}  // namespace ARW_NS

using namespace ARW_NS;

extern "C" int32_t ARW_API(health_gate)(const arw_health_field* fields,
                             int32_t nfields, uint64_t chunk,
                             uint32_t nchunks, uint64_t* result,
                             void* stream)
{
    if (!fields || !result || nfields < 0 || nfields > 65535 || chunk == 0
        || nchunks == 0)
        return ARW_E_ARG;
    // the record starts at its identity; then one launch over every
    // (chunk, field) pair
    hipStream_t st = (hipStream_t)stream;
    auto* res = reinterpret_cast<unsigned long long*>(result);
    hipLaunchKernelGGL(k_health_init, dim3(1), dim3(64), 0, st, res);
    if (nfields > 0)
        hipLaunchKernelGGL(k_health_gate, dim3(nchunks, (unsigned)nfields),
                           dim3(256), 0, st, fields, (int)nfields,
                           (unsigned long long)chunk, res);
    return hipGetLastError() == hipSuccess ? ARW_OK : ARW_E_LAUNCH;
}
