// pfdd_plane.hip -- the two-pass ("plane") form of the GL iteration for grids
// whose y-z plane fits in the 160 KB LDS of a gfx950 CU (FP32 up to 128 x 128,
// FP64 up to 80 x 80).  The five-kernel iteration (fwd z, fwd y, x + B, inv y,
// inv z + update) becomes three:
//
//   A  k_plane_fwd  xi[a][i][.][.] -> T[a][i][.][.]: 2-D FFT of each y-z plane
//                   (along k, then along j), written back in the same layout
//   B  x pass       T in place along i (k_mid_otf in its plane layout)
//   C  k_plane_upd  T + xi_old -> xi_new: inverse 2-D FFT of each plane (along
//                   j, then along k) fused with the GL update and its sums
//
// so an iteration moves 7 S N e bytes instead of 11 S N e.  Each line is
// transformed with the radix plan and twiddles of the line passes
// (fft_stages_g), and the update is the same gl_update_cell, so the field is
// bit-identical to the five-kernel iteration's.
//
// Both kernels are persistent: a block walks planes blockIdx.x, + gridDim.x,
// ... and loads the next plane into registers while the current one is
// transformed (and, in C, updated), so HBM traffic and the FFT arithmetic
// overlap even at one resident block per CU.  Twiddles are staged in LDS: global twiddle
// loads would share the in-order vmcnt counter with the prefetch, and the
// first twiddle wait would also wait for the next plane.
//
// Reference functions replaced (fft_fftw_slab.cpp): forward_mode1 :1953-1964
// (first two axes), backward_mode1 :2012-2023 (last two axes),
// core_energy_* :1130-1418 and update_order_parameter :1746-1790.
#include "pfdd_kern.hpp"

namespace pfdd {

// Threads per block of a plane kernel: one thread per 8-16 elements (FP32)
// or 9-18 (FP64, whose 16-byte prefetch buffers need twice the registers).
template <typename R, int NY, int NZ> constexpr int plane_threads()
{
    return (NY * NZ >= 8192 && sizeof(R) == 4) ? 1024 : NY * NZ >= 4096 ? 512 : 256;
}

// Tile of one plane: element (j, k) at k * (NY + 1) + j.
template <typename R, int NY, int NZ> constexpr long long plane_lds_bytes()
{
    return (long long)NZ * (NY + 1) * 2 * sizeof(R) + (long long)(NY + NZ) * 2 * sizeof(R);
}

// The plane must fit in the 160 KB of LDS; FP64 planes are taken up to 80 x
// 80 (at 96 x 96 the FP64 kernels exceed the register file and spill).
template <typename R, int NY, int NZ> constexpr bool plane_fits()
{
    return plane_lds_bytes<R, NY, NZ>() <= 160 * 1024 && (sizeof(R) == 4 || NY * NZ <= 80 * 80);
}

template <int TOT, int NT, typename V>
__device__ __forceinline__ void plane_fetch(V *r, const V *__restrict__ src, int tid)
{
    constexpr int PER = (TOT + NT - 1) / NT;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = tid + t * NT;
        if (TOT % NT == 0 || e < TOT) r[t] = ld_stream(&src[e]);
    }
}

template <int NY, int NZ, int NT, typename V>
__device__ __forceinline__ void plane_put(const Tile<V> lds, const V *r, int tid)
{
    constexpr int TOT = NY * NZ, PER = (TOT + NT - 1) / NT, LDY = NY + 1;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = tid + t * NT;
        if (TOT % NT == 0 || e < TOT) lds.st((e % NZ) * LDY + e / NZ, r[t]);
    }
}

template <int NY, int NZ, int NT, typename V>
__device__ __forceinline__ void plane_twiddles(V *tws, const V *__restrict__ twy, const V *__restrict__ twz, int tid)
{
    for (int m = tid; m < NY; m += NT) tws[m] = twy[m];
    for (int m = tid; m < NZ; m += NT) tws[NY + m] = twz[m];
}

// A: forward 2-D FFT of every y-z plane (plane index pl = a * nx + i).
template <typename R, int NY, int NZ>
__global__ void __launch_bounds__((plane_threads<R, NY, NZ>()))
    k_plane_fwd(const typename cvec<R>::type *__restrict__ in, typename cvec<R>::type *__restrict__ out,
                long long nplanes, const typename cvec<R>::type *__restrict__ twy,
                const typename cvec<R>::type *__restrict__ twz, const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NT = plane_threads<R, NY, NZ>();
    constexpr int LDY = NY + 1, TOT = NY * NZ, PER = (TOT + NT - 1) / NT;
    PFDD_TILE(V, lds, NZ * LDY);
    __shared__ V tws[NY + NZ];
    const int tid = threadIdx.x;
    plane_twiddles<NY, NZ, NT>(tws, twy, twz, tid);
    long long pl = blockIdx.x;
    V r[PER];
    if (pl < nplanes) plane_fetch<TOT, NT>(r, in + pl * TOT, tid);
    for (; pl < nplanes; pl += gridDim.x) {
        plane_put<NY, NZ, NT>(lds, r, tid);
        __syncthreads();
        const long long pn = pl + gridDim.x;
        if (pn < nplanes) plane_fetch<TOT, NT>(r, in + pn * TOT, tid);
        // along k: lines j (unit stride), positions k (stride NY + 1)
        fft_stages_g<NZ, 1, NY, 1, LDY, NT, false>(lds, tws + NY, tid);
        // along j: lines k (stride NY + 1), positions j (unit stride)
        fft_stages_g<NY, 1, NZ, LDY, 1, NT, false>(lds, tws, tid);
        V *o = out + pl * TOT;
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int e = tid + t * NT;
            if (TOT % NT == 0 || e < TOT) st_stream(&o[e], lds.ld((e % NZ) * LDY + e / NZ));
        }
        __syncthreads();
    }
}

// C: inverse 2-D FFT of every plane of T fused with the GL update of the
// plane's cells (gl_update_cell), and per-plane sums of xinorm^2, xiave, the
// core energy and the stress work (partial[slot * pstride + plane]).  A plane holds one system, so tau and the core
// coefficients are uniform over it.
template <typename R, int NY, int NZ, bool GEN>
__global__ void __launch_bounds__((plane_threads<R, NY, NZ>()))
    k_plane_upd(const typename cvec<R>::type *__restrict__ T, const typename cvec<R>::type *__restrict__ xo,
                typename cvec<R>::type *__restrict__ xn, const uint8_t *__restrict__ mask,
                const long long *__restrict__ prev, const R *__restrict__ amp, long long nplanes, int nx,
                StepParams P, const typename cvec<R>::type *__restrict__ twy,
                const typename cvec<R>::type *__restrict__ twz, double *__restrict__ partial, long long pstride,
                const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NT = plane_threads<R, NY, NZ>();
    constexpr int LDY = NY + 1, TOT = NY * NZ, PER = (TOT + NT - 1) / NT;
    PFDD_TILE(V, lds, NZ * LDY);
    __shared__ V tws[NY + NZ];
    __shared__ double red[4 * (NT / 64)];
    // the plane's core coefficients, staged in LDS: the sine series indexes
    // them with a run-time loop counter
    __shared__ R lcf[MAXM];
    const int tid = threadIdx.x;
    plane_twiddles<NY, NZ, NT>(tws, twy, twz, tid);
    const R cddt = (R)P.cddt, invN = (R)P.invN;
    const int nm = P.ncoef;
    // The next plane of T is in flight in registers during the transform and
    // the update of the current one; the old field is read by the update loop
    // itself, a few cells per thread at a time (holding it in registers
    // across the transform as well does not fit the register budget of
    // 1024-thread blocks).
    constexpr int UNR = GEN ? 1 : 4;
    long long pl = blockIdx.x;
    V r[PER];
    if (pl < nplanes) plane_fetch<TOT, NT>(r, T + pl * TOT, tid);
    for (; pl < nplanes; pl += gridDim.x) {
        const int a = (int)(pl / nx); // system in the batch
        const int as = a % P.S;       // system within its member
        if (tid < MAXM) lcf[tid] = (R)P.coef[as * MAXM + tid];
        plane_put<NY, NZ, NT>(lds, r, tid);
        __syncthreads();
        const long long pn = pl + gridDim.x;
        if (pn < nplanes) plane_fetch<TOT, NT>(r, T + pn * TOT, tid);
        fft_stages_g<NY, 1, NZ, LDY, 1, NT, true>(lds, tws, tid);
        fft_stages_g<NZ, 1, NY, 1, LDY, NT, true>(lds, tws + NY, tid);
        const long long base = pl * TOT;
        const R tau = (R)(P.taud ? P.taud[a] : P.tau[a]);
        // plain FP64 sums over the thread's cells of the plane (at most 16),
        // the block tree of block_sum_comp, one partial per plane
        double p_norm = 0.0, p_ave = 0.0, p_core = 0.0, p_str = 0.0;
#pragma unroll UNR
        for (int e = tid; e < TOT; e += NT) {
            const long long idx = base + e;
            const V g = lds.ld((e % NZ) * LDY + e / NZ);
            const V x = ld_stream(&xo[idx]);
            R dens;
            double dn;
            const V y = gl_update_cell<R, GEN>(g, x, idx, tau, as % 3, [&](int m) { return lcf[m]; }, P, nm,
                                               cddt, invN, mask, prev, xo, amp, dn, dens);
            st_stream(&xn[idx], y);
            p_norm += dn;
            p_ave += (double)y.x;
            p_core += (double)dens;
            p_str += (double)tau * (double)x.x;
        }
        double v[4] = {p_norm, p_ave, p_core, p_str};
        block_sum_comp<4, NT>(v, red, tid);
        if (tid == 0) {
            partial[0 * pstride + pl] = v[0];
            partial[1 * pstride + pl] = v[1];
            partial[2 * pstride + pl] = v[2];
            partial[3 * pstride + pl] = v[3];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- launchers
#define PFDD_PLANE_SIZES(X) X(16) X(32) X(48) X(64) X(80) X(96) X(128)

// Resident blocks per CU of kernel `which` (0 forward, 1 update, 2 update of
// the coupled / Fourier core models) for N x N planes.
template <typename R, int N> int plane_occupancy_n(int which)
{
    auto occ = [](const void *f) {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, f, plane_threads<R, N, N>(), 0) != hipSuccess) o = 1;
        return o > 0 ? o : 1;
    };
    static const int o[3] = {occ((const void *)k_plane_fwd<R, N, N>), occ((const void *)k_plane_upd<R, N, N, false>),
                             occ((const void *)k_plane_upd<R, N, N, true>)};
    return o[which < 0 ? 0 : which > 2 ? 2 : which];
}

template <typename R> bool plane_supported(int ny, int nz)
{
    if (ny != nz) return false;
#define PFDD_CASE(n_) if (ny == n_) return plane_fits<R, n_, n_>();
    PFDD_PLANE_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return false;
}

template <typename R> int plane_blocks(int ny, int nz, long long nplanes, int ncu, int which)
{
    int occ = 1;
    if (ny == nz) {
#define PFDD_CASE(n_)                                                               \
    if (ny == n_) {                                                                 \
        if constexpr (plane_fits<R, n_, n_>()) occ = plane_occupancy_n<R, n_>(which); \
    }
        PFDD_PLANE_SIZES(PFDD_CASE)
#undef PFDD_CASE
    }
    const long long cap = (long long)ncu * occ;
    return (int)(nplanes < cap ? nplanes : cap);
}

template <typename R>
hipError_t launch_plane_fwd(int ny, int nz, const void *in, void *out, long long nplanes, int nblocks,
                            const void *twy, const void *twz, const int *done, hipStream_t st)
{
    typedef typename cvec<R>::type V;
    if (ny != nz) return hipErrorInvalidValue;
#define PFDD_CASE(n_)                                                                                   \
    if (ny == n_) {                                                                                     \
        if constexpr (plane_fits<R, n_, n_>()) {                                                        \
            PFDD_LAUNCH((k_plane_fwd<R, n_, n_>), dim3((unsigned)nblocks), dim3(plane_threads<R, n_, n_>()), \
                        st, (const V *)in, (V *)out, nplanes, (const V *)twy, (const V *)twz, done);    \
            return hipGetLastError();                                                                   \
        }                                                                                               \
    }
    PFDD_PLANE_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_plane_upd(int ny, int nz, const void *T, const void *xo, void *xn, const uint8_t *mask,
                            const long long *prev, const void *amp, long long nplanes, int nx, int nblocks,
                            const StepParams &P, const void *twy, const void *twz, double *partial,
                            long long pstride, const int *done, hipStream_t st)
{
    typedef typename cvec<R>::type V;
    if (ny != nz) return hipErrorInvalidValue;
    const bool gen = P.model == PFDD_CORE_EXTENDED || P.model == PFDD_CORE_PYRII || P.model == PFDD_CORE_FOURIER;
#define PFDD_CASE(n_)                                                                                        \
    if (ny == n_) {                                                                                          \
        if constexpr (plane_fits<R, n_, n_>()) {                                                             \
            if (gen)                                                                                         \
                PFDD_LAUNCH((k_plane_upd<R, n_, n_, true>), dim3((unsigned)nblocks),                         \
                            dim3(plane_threads<R, n_, n_>()), st, (const V *)T, (const V *)xo, (V *)xn, mask,   \
                            prev, (const R *)amp, nplanes, nx, P, (const V *)twy, (const V *)twz, partial,   \
                            pstride, done);                                                                  \
            else                                                                                             \
                PFDD_LAUNCH((k_plane_upd<R, n_, n_, false>), dim3((unsigned)nblocks),                        \
                            dim3(plane_threads<R, n_, n_>()), st, (const V *)T, (const V *)xo, (V *)xn, mask,   \
                            prev, (const R *)amp, nplanes, nx, P, (const V *)twy, (const V *)twz, partial,   \
                            pstride, done);                                                                  \
            return hipGetLastError();                                                                        \
        }                                                                                                    \
    }
    PFDD_PLANE_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

#define PFDD_INST(R)                                                                                           \
    template bool plane_supported<R>(int, int);                                                                \
    template int plane_blocks<R>(int, int, long long, int, int);                                              \
    template hipError_t launch_plane_fwd<R>(int, int, const void *, void *, long long, int, const void *,        \
                                            const void *, const int *, hipStream_t);                           \
    template hipError_t launch_plane_upd<R>(int, int, const void *, const void *, void *, const uint8_t *,     \
                                            const long long *, const void *, long long, int, int,              \
                                            const StepParams &, const void *, const void *, double *, long long, \
                                            const int *, hipStream_t);
PFDD_INST(double)
PFDD_INST(float)
#undef PFDD_INST

} // namespace pfdd
