// pfdd_update.hip -- the last inverse pass fused with the GL update, and its
// reverse-mode counterpart, with their launchers.
#include "pfdd_kern.hpp"

namespace pfdd {

// ------------------------------------- pass 5: inverse z FFT + fused update
// update_order_parameter (fft_fftw_slab.cpp:1746-1790) with the core-energy
// derivative of core_energy_* evaluated on the old field, as the reference
// does (it reads data_fftw = xi before the update).  Masked cells keep their
// value; their xinorm term uses the previous unmasked cell's old value in
// (a, i, j, k) order (prev[] precomputed on the host, -1 = the initial 0).
// The FP64 sine-series update on its smallest tile at 32-128 points (8
// lines at 128 points, 19-20 KB of LDS: 8 blocks per CU) asks for 8 waves per
// SIMD: 64 instead of 68 VGPRs, 229 -> 224.5 us at 128^3 x 12.  The same
// hint made the FP32 one 9 % slower.
template <typename R, int N, int TQ, bool GEN> constexpr int update_waves()
{
    return (sizeof(R) == 8 && TQ == 0 && !GEN && N >= 32 && N <= 128) ? 8 : 1;
}

template <typename R, int N, int TQ, bool GEN>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(update_waves<R, N, TQ, GEN>(), 8)))
    k_update(const typename cvec<R>::type *__restrict__ T1, const typename cvec<R>::type *__restrict__ xo,
             typename cvec<R>::type *__restrict__ xn, const uint8_t *__restrict__ mask,
             const long long *__restrict__ prev, const R *__restrict__ amp, LineMap mi, LineMap mo,
             long long nlines, long long lines_per_sys, StepParams P,
             const typename cvec<R>::type *__restrict__ tw, double *__restrict__ partial,
             long long pstride, const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NL = tile_lines_u<R, N, TQ>(), LDP = NL + 1, TOT = N * NL;
    PFDD_TILE(V, lds, N * LDP);
    __shared__ long long bin[NL], bout[NL];
    // per-line tau and core coefficients: read once per block here, since
    // kernel-argument arrays indexed per lane compile to per-lane memory loads
    __shared__ R ltau[NL], lcoef[MAXM * NL];
    __shared__ int lsub[NL]; // position of the line's system in its plane triple (coupled model)
    __shared__ double red[4 * (NTHR / 64)];
    const int tid = threadIdx.x;
    const long long L0 = (long long)blockIdx.x * NL;
    for (int l = tid; l < NL; l += NTHR) {
        const long long L = L0 + l;
        const bool ok = L < nlines;
        bin[l] = ok ? line_base(mi, L) : -1;
        bout[l] = ok ? line_base(mo, L) : -1;
        const int a = ok ? (int)(L / lines_per_sys) : 0; // system in the batch
        const int as = a % P.S;                           // system within its member
        ltau[l] = (R)(P.taud ? P.taud[a] : P.tau[a]);
        for (int m = 0; m < MAXM; ++m) lcoef[m * NL + l] = (R)P.coef[as * MAXM + m];
        lsub[l] = as % 3;
    }
    __syncthreads();
    tile_load<N, NL, NTHR>(lds, T1, bin, mi.SP, tid);
    // the old field of this thread's output cells, all loads in flight at
    // once: 8-byte elements while the FFT runs, 16-byte elements right after
    // it (holding them in registers across the FFT measured slower)
    constexpr int PER = (TOT + NTHR - 1) / NTHR;
    constexpr bool PREF = sizeof(V) < 16;
    V xold[PER];
    if constexpr (PREF) {
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int e = tid + t * NTHR;
            const long long b = (TOT % NTHR == 0 || e < TOT) ? bout[e / N] : -1;
            xold[t] = b >= 0 ? ld_stream(&xo[b + e % N]) : mk<V>((R)0, (R)0);
        }
    }
    __syncthreads();
    fft_stages<N, 1, NL, NTHR, true>(lds, tw, tid);
    if constexpr (!PREF) {
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int e = tid + t * NTHR;
            const long long b = (TOT % NTHR == 0 || e < TOT) ? bout[e / N] : -1;
            xold[t] = b >= 0 ? ld_stream(&xo[b + e % N]) : mk<V>((R)0, (R)0);
        }
    }

    const R cddt = (R)P.cddt, invN = (R)P.invN;
    const int nm = P.ncoef;
    // each thread sums at most 16 cells in FP64 (relative error below 16 u);
    // the block level is a pairwise tree with a compensated wave combine
    // (block_sum_comp), the grid level a Kahan sum (k_reduce)
    double s_norm = 0.0, s_ave = 0.0, s_core = 0.0, s_str = 0.0;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = tid + t * NTHR;
        if (TOT % NTHR != 0 && e >= TOT) continue;
        const int l = e / N, p = e % N; // mo.SP == 1: rows of xi are contiguous
        const long long b = bout[l];
        if (b < 0) continue;
        const long long idx = b + p;
        const V g = lds.ld(p * LDP + l);
        const V x = xold[t];
        const R tau = ltau[l];
        R dens;
        double dn;
        const V y = gl_update_cell<R, GEN>(g, x, idx, tau, lsub[l], [&](int m) { return lcoef[m * NL + l]; }, P, nm,
                                           cddt, invN, mask, prev, xo, amp, dn, dens);
        st_stream(&xn[idx], y);
        s_norm += dn;
        s_ave += (double)y.x;
        s_core += (double)dens;
        s_str += (double)tau * (double)x.x;
    }
    double v[4] = {s_norm, s_ave, s_core, s_str};
    block_sum_comp<4, NTHR>(v, red, tid);
    if (tid == 0) {
        partial[0 * pstride + blockIdx.x] = v[0];
        partial[1 * pstride + blockIdx.x] = v[1];
        partial[2 * pstride + blockIdx.x] = v[2];
        partial[3 * pstride + blockIdx.x] = v[3];
    }
}

// ------------------------------------------ adjoint of pass 5 (reverse mode)
// One GL iteration (k_update) maps xi to
//   xi' = xi - cddt U (g/N - tau + dE(Re xi))      (Re part; Im: xi' = xi - cddt U g/N)
// with U = 1 on unmasked cells and g = IFFT(B FFT xi).  The elastic operator
// is self-adjoint (B(k) real, symmetric in (a, b) and even in k), so for
// lam' = dL/dxi' (Re and Im taken as independent real variables)
//   mu  = U lam'
//   lam = lam' - cddt (IFFT(B FFT mu)/N + dE'(Re xi) Re mu)   (dE' acts on Re only)
//   dL/dtau_a   += cddt sum_r Re mu_a
//   dL/dc_{a,m} += -cddt sum_r Re mu_a A m pi sin(2 m pi Re xi_a)
// with dE = A sum_m c_m m pi sin(2 m pi x), dE' = A sum_m c_m 2 m^2 pi^2 cos(2 m pi x).
// The extended model couples the three systems of a plane: there dE'(Re xi)
// Re mu becomes the Hessian of the plane density applied to the triple's
// Re mu (core_extended_adj).
// T1 holds mu after passes 1-4; this kernel applies the last inverse pass and
// the update above, and leaves per-block sums of Re mu (slot 0) and of
// Re mu A m pi sin(2 m pi x) (slot m) in gpart.  Every block covers lines of
// one system (the host checks nx*ny % NL == 0).
template <typename R, int N, int TQ>
__global__ void __launch_bounds__(NTHR)
    k_adj_update(const typename cvec<R>::type *__restrict__ T1, const typename cvec<R>::type *__restrict__ lam,
                 const typename cvec<R>::type *__restrict__ xs, typename cvec<R>::type *__restrict__ lam_out,
                 const uint8_t *__restrict__ mask, const R *__restrict__ amp, LineMap mi, LineMap mo,
                 long long nlines, long long lines_per_sys, StepParams P,
                 const typename cvec<R>::type *__restrict__ tw, double *__restrict__ gpart, long long gstride,
                 const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NL = tile_lines_u<R, N, TQ>(), LDP = NL + 1, TOT = N * NL;
    constexpr int PER = (TOT + NTHR - 1) / NTHR;
    constexpr int NG = 1 + MAXM;
    PFDD_TILE(V, lds, N * LDP);
    __shared__ long long bin[NL], bout[NL];
    __shared__ double red[NG * (NTHR / 64)];
    const int tid = threadIdx.x;
    const long long L0 = (long long)blockIdx.x * NL;
    for (int l = tid; l < NL; l += NTHR) {
        const long long L = L0 + l;
        const bool ok = L < nlines;
        bin[l] = ok ? line_base(mi, L) : -1;
        bout[l] = ok ? line_base(mo, L) : -1;
    }
    __syncthreads();
    tile_load<N, NL, NTHR>(lds, T1, bin, mi.SP, tid);
    __syncthreads();
    fft_stages<N, 1, NL, NTHR, true>(lds, tw, tid);

    const int a = (int)(L0 / lines_per_sys) % P.S;
    const R cddt = (R)P.cddt, invN = (R)P.invN;
    const int nm = P.ncoef;
    double acc[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) acc[q] = 0.0;
#pragma unroll 2
    for (int t = 0; t < PER; ++t) {
        const int e = tid + t * NTHR;
        if (TOT % NTHR != 0 && e >= TOT) continue;
        const int l = e / N, p = e % N;
        const long long b = bout[l];
        if (b < 0) continue;
        const long long idx = b + p;
        const V g = lds.ld(p * LDP + l);
        const V lm = lam[idx];
        const V x = xs[idx];
        const R mr = (mask && mask[idx]) ? (R)0 : lm.x;
        const R A = amp ? (R)amp[idx] : (R)1;
        // sin, cos of 2 pi m x by the angle-addition recurrence from m = 1
        R sp, cp;
        sincospi_poly(x.x, sp, cp);
        const R s1 = (R)2 * sp * cp, c1 = cp * cp - sp * sp;
        R sm = s1, cm = c1, d2 = (R)0, hv = (R)0;
        if (P.model == PFDD_CORE_EXTENDED) {
            // the triple of systems of this plane at the same cell (no amp[])
            const int j = a % 3;
            const long long b0 = idx - (long long)j * P.ncell;
            R y[3], m3[3], cc[7];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const long long iq = b0 + q * P.ncell;
                y[q] = q == j ? x.x : xs[iq].x;
                m3[q] = q == j ? mr : ((mask && mask[iq]) ? (R)0 : lam[iq].x);
            }
#pragma unroll
            for (int q = 0; q < 7; ++q) cc[q] = (R)P.coef[a * MAXM + q];
            double gc[7];
            core_extended_adj<R>(y, m3, j, cc, hv, gc);
#pragma unroll
            for (int q = 0; q < 7; ++q) acc[1 + q] += gc[q];
        } else if (P.model == PFDD_CORE_FOURIER || P.model == PFDD_CORE_PYRII) {
            // Fourier density a0 + sum_m a_m cos(w x) + b_m sin(w x), w = 2 pi m:
            //   dE = sum_m w (b_m cos - a_m sin),  dE' = -sum_m w^2 (a_m cos + b_m sin)
            //   d(dE)/d a_m = -w sin(w x),  d(dE)/d b_m = w cos(w x)   (slots m, 4 + m)
            // pyrII as executed: only the last cell of each system carries dE
            if (P.model == PFDD_CORE_FOURIER || (idx % P.ncell) == P.ncell - 1) {
#pragma unroll
                for (int m = 1; m <= 4; ++m) {
                    const R am = (R)P.coef[a * MAXM + m - 1], bm = (R)P.coef[a * MAXM + m + 3];
                    const double w = 2.0 * M_PI * m;
                    d2 -= (R)(w * w) * (am * cm + bm * sm);
                    acc[m] += (double)mr * (double)A * (-w * (double)sm);
                    acc[4 + m] += (double)mr * (double)A * (w * (double)cm);
                    const R sn = sm * c1 + cm * s1, cn = cm * c1 - sm * s1;
                    sm = sn;
                    cm = cn;
                }
            }
        } else {
#pragma unroll
            for (int m = 1; m <= MAXM; ++m) {
                if (m <= nm) {
                    const R c = (R)P.coef[a * MAXM + m - 1];
                    d2 += c * ((R)(2 * m * m) * (R)(M_PI * M_PI)) * cm;
                    acc[m] += (double)mr * (double)A * (double)m * M_PI * (double)sm;
                    const R sn = sm * c1 + cm * s1, cn = cm * c1 - sm * s1;
                    sm = sn;
                    cm = cn;
                }
            }
        }
        V y;
        y.x = lm.x - cddt * (g.x * invN + A * d2 * mr + hv);
        y.y = lm.y - cddt * (g.y * invN);
        lam_out[idx] = y;
        acc[0] += (double)mr;
    }
    block_sum<NG, NTHR>(acc, red, tid);
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NG; ++q) gpart[q * gstride + blockIdx.x] = acc[q];
    }
}


template <typename R>
hipError_t launch_update(int n, int tq, const void *T1, const void *xo, void *xn, const uint8_t *mask,
                         const long long *prev, const void *amp, LineMap mi, LineMap mo, long long nlines,
                         long long lps, const StepParams &P, const void *tw, double *partial,
                         long long pstride, const int *done, hipStream_t st)
{
    typedef typename cvec<R>::type V;
    // the coupled and Fourier core models get their own instantiation (one
    // tile shape) so that the common sine-series kernel carries none of
    // their registers
    const bool gen = P.model == PFDD_CORE_EXTENDED || P.model == PFDD_CORE_PYRII || P.model == PFDD_CORE_FOURIER;
#define PFDD_CASE(n_)                                                                                \
    if (n == n_) {                                                                                   \
        if (gen) {                                                                                   \
            constexpr int NL = tile_lines_u<R, n_, 0>();                                             \
            const unsigned nb = (unsigned)((nlines + NL - 1) / NL);                                  \
            PFDD_LAUNCH((k_update<R, n_, 0, true>), dim3(nb), dim3(NTHR), st, (const V *)T1,         \
                        (const V *)xo, (V *)xn, mask, prev, (const R *)amp, mi, mo, nlines, lps, P,  \
                        (const V *)tw, partial, pstride, done);                                      \
            return hipGetLastError();                                                                \
        }                                                                                            \
        PFDD_TQ_DISPATCH(tq, {                                                                       \
            constexpr int NL = tile_lines_u<R, n_, TQ>();                                            \
            const unsigned nb = (unsigned)((nlines + NL - 1) / NL);                                  \
            PFDD_LAUNCH((k_update<R, n_, TQ, false>), dim3(nb), dim3(NTHR), st, (const V *)T1,       \
                               (const V *)xo, (V *)xn, mask, prev, (const R *)amp, mi, mo, nlines, lps, \
                               P, (const V *)tw, partial, pstride, done);                            \
        })                                                                                           \
        return hipGetLastError();                                                                    \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_adj_update(int n, int tq, const void *T1, const void *lam, const void *xs, void *lam_out,
                             const uint8_t *mask, const void *amp, LineMap mi, LineMap mo, long long nlines,
                             long long lps, const StepParams &P, const void *tw, double *gpart,
                             long long gstride, const int *done, hipStream_t st)
{
    typedef typename cvec<R>::type V;
#define PFDD_CASE(n_)                                                                                  \
    if (n == n_) {                                                                                     \
        PFDD_TQ_DISPATCH(tq, {                                                                         \
            constexpr int NL = tile_lines_u<R, n_, TQ>();                                              \
            const unsigned nb = (unsigned)((nlines + NL - 1) / NL);                                    \
            hipLaunchKernelGGL((k_adj_update<R, n_, TQ>), dim3(nb), dim3(NTHR), 0, st, (const V *)T1,  \
                               (const V *)lam, (const V *)xs, (V *)lam_out, mask, (const R *)amp, mi, mo, \
                               nlines, lps, P, (const V *)tw, gpart, gstride, done);                   \
        })                                                                                             \
        return hipGetLastError();                                                                      \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

#define PFDD_INST(R)                                                                                          \
    template hipError_t launch_update<R>(int, int, const void *, const void *, void *, const uint8_t *,          \
                                         const long long *, const void *, LineMap, LineMap, long long, long long, \
                                         const StepParams &, const void *, double *, long long, const int *,      \
                                         hipStream_t);                                                            \
    template hipError_t launch_adj_update<R>(int, int, const void *, const void *, const void *, void *,         \
                                             const uint8_t *, const void *, LineMap, LineMap, long long,          \
                                             long long, const StepParams &, const void *, double *, long long,    \
                                             const int *, hipStream_t);
PFDD_INST(double)
PFDD_INST(float)
#undef PFDD_INST

} // namespace pfdd
