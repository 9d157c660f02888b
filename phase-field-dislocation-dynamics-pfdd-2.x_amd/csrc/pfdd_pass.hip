// pfdd_pass.hip -- the line-pass kernel (FFT along one axis of every line of
// a batch, transposed write through LDS) and its launcher.
#include "pfdd_kern.hpp"

namespace pfdd {

// ------------------------------------------------- generic FFT pass (1, 2, 4)
// TQ: tile shape (tile_lines_q); PL: radix plan (0 radix 8 first, 1 radix 16
// where it saves a stage).
template <typename R, int N, bool INV, int TQ, int PL>
__global__ void __launch_bounds__(NTHR)
    k_fft_pass(const typename cvec<R>::type *__restrict__ in, typename cvec<R>::type *__restrict__ out,
               LineMap mi, LineMap mo, long long nlines, const typename cvec<R>::type *__restrict__ tw,
               const int *__restrict__ done, const uint8_t *__restrict__ zmask)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NL = tile_lines_q<R, N, TQ>(), LDP = NL + 1, TOT = N * NL;
    PFDD_TILE(V, lds, N * LDP);
    __shared__ long long bin[NL], bout[NL];
    const int tid = threadIdx.x;
    const long long L0 = (long long)blockIdx.x * NL;
    for (int l = tid; l < NL; l += NTHR) {
        const long long L = L0 + l;
        const bool ok = L < nlines;
        bin[l] = ok ? line_base(mi, L) : -1;
        bout[l] = ok ? line_base(mo, L) : -1;
    }
    __syncthreads();
    tile_load<N, NL, NTHR>(lds, in, bin, mi.SP, tid);
    __syncthreads();
    if (zmask) {
        // input elements whose zmask byte is set enter the transform as 0
        for (int e = tid; e < TOT; e += NTHR) {
            const int l = e / N, p = e % N;
            const long long b = bin[l];
            if (b >= 0 && zmask[b + p * mi.SP]) lds.st(p * LDP + l, mk<V>((R)0, (R)0));
        }
        __syncthreads();
    }
    if constexpr (PL == 1) fft_stages_x<N, 1, NL, NTHR, INV>(lds, tw, tid);
    else fft_stages<N, 1, NL, NTHR, INV>(lds, tw, tid);
    tile_store<N, NL, NTHR>(out, lds, bout, mo.SP, tid);
}


template <typename R, bool INV>
hipError_t launch_pass(int n, const void *in, void *out, LineMap mi, LineMap mo, long long nlines,
                       const void *tw, const int *done, hipStream_t st, const uint8_t *zmask,
                       int tq)
{
    typedef typename cvec<R>::type V;
    // tq = shape + 3 * plan; plan 1 falls back to plan 0 where they coincide
    const int shape = tq % 3, plan = tq / 3;
#define PFDD_CASE(n_)                                                                               \
    if (n == n_) {                                                                                  \
        PFDD_TQ_DISPATCH(shape, {                                                                   \
            constexpr int NL = tile_lines_q<R, n_, TQ>();                                           \
            const unsigned nb = (unsigned)((nlines + NL - 1) / NL);                                 \
            if (plan == 1 && radix16_plan_differs<n_>())                                            \
                PFDD_LAUNCH((k_fft_pass<R, n_, INV, TQ, radix16_plan_differs<n_>() ? 1 : 0>),      \
                            dim3(nb), dim3(NTHR), st, (const V *)in, (V *)out, mi, mo, nlines,      \
                            (const V *)tw, done, zmask);                                            \
            else                                                                                    \
                PFDD_LAUNCH((k_fft_pass<R, n_, INV, TQ, 0>), dim3(nb), dim3(NTHR), st,              \
                            (const V *)in, (V *)out, mi, mo, nlines, (const V *)tw, done, zmask);   \
        })                                                                                          \
        return hipGetLastError();                                                                   \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

template hipError_t launch_pass<double, false>(int, const void *, void *, LineMap, LineMap, long long, const void *,
                                               const int *, hipStream_t, const uint8_t *, int);
template hipError_t launch_pass<double, true>(int, const void *, void *, LineMap, LineMap, long long, const void *,
                                              const int *, hipStream_t, const uint8_t *, int);
template hipError_t launch_pass<float, false>(int, const void *, void *, LineMap, LineMap, long long, const void *,
                                              const int *, hipStream_t, const uint8_t *, int);
template hipError_t launch_pass<float, true>(int, const void *, void *, LineMap, LineMap, long long, const void *,
                                             const int *, hipStream_t, const uint8_t *, int);

} // namespace pfdd
