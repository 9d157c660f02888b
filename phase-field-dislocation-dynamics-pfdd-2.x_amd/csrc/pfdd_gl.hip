// pfdd_gl.hip -- gfx950 implementation of the PFDD Ginzburg-Landau step and
// its C ABI (include/pfdd_gl.h).
//
// One GL iteration of SolveGL::iterate (solve_GL.cpp:86-148) is five kernels
// over the S order parameters of an nx*ny*nz grid, with the field kept in
// three device buffers (two ping-pong copies of xi and one temporary T1):
//
//   1  fwd z   xi[a][i][j][k]  -> T1[a][i][k][j]   FFT along k, transposed write
//   2  fwd y   T1[a][i][k][j]  -> T2[a][k][j][i]   FFT along j, transposed write
//   3  x+B     T2 in place: FFT along i, ghat_a = sum_b B_ab xihat_b (packed
//              symmetric B, FP64 or FP32), inverse FFT along i, plus the
//              elastic-energy partial sums
//   4  inv y   T2 -> T1[a][i][k][j]
//   5  inv z   T1 -> g, fused core-energy derivative, applied stress, explicit
//              Euler update, obstacle mask and reductions: xi_old -> xi_new
//   +  one single-block kernel reduces the per-block partials in fixed order.
//
// T2 is the ping-pong copy of xi that is not current, so the whole iteration
// needs 3*S*N complex values plus B.  Every global access is a run of at least
// 8 consecutive complex values (the pass transposes through LDS), and every
// kernel reads and writes each field element exactly once.
//
// Reference functions replaced (fft_fftw_slab.cpp): core_energy_* :1130-1418,
// forward_mode1 :1953-1964, internal_energy :1670-1695, backward_mode1
// :2012-2023, update_order_parameter :1746-1790, prepare_next_itr :1838-1862,
// init_loop :1628-1666 (no longer needed), Bmatrix :466-638, frec :280-367.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdarg>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pfdd_fft.hpp"
#include "../../include/pfdd_gl.h"

using namespace pfdd;

namespace {

constexpr int MAXS = PFDD_GL_MAX_SYSTEMS;
constexpr int MAXM = PFDD_GL_MAX_COEF;
constexpr int NTHR = 256;
constexpr int NSLOT = 5; // xinorm^2, xiave, e_core, e_stress, e_el
constexpr int NRES = 8;  // xinorm, xiave, e_el, e_core, e_stress, e_total, e_core_ref, nonfinite

thread_local std::string g_err;

// Kernel-timing events for the next launch of the GL iteration's kernels
// (set by the profiler around one launch at a time): the dispatch itself
// records them (hipExtLaunchKernelGGL), so they bracket exactly the kernel.
thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;

#define PFDD_LAUNCH(kern, grid, block, st, ...)                                                        \
    do {                                                                                              \
        if (t_ev_start)                                                                               \
            hipExtLaunchKernelGGL(kern, grid, block, 0, st, t_ev_start, t_ev_stop, 0, __VA_ARGS__);    \
        else                                                                                          \
            hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__);                                \
    } while (0)

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(PFDD_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

struct StepParams {
    double tau[MAXS];          // resolved shear per system (single-run handles)
    const double *taud;        // device tau per system of a batch (nullptr: use tau[])
    double coef[MAXS * MAXM];  // core coefficients per system of one member
    double a0;                 // constant term of the Fourier core models
    double cddt;
    double invN;
    long long ncell;           // cells per system
    int model;                 // PFDD_CORE_*
    int ncoef;
    int S;                     // systems per member
};

// Coupled gamma surface of the three systems of one plane (core_energy_extended,
// fft_fftw_slab.cpp:1243-1354): density f(y1, y2, y3) and its derivative with
// respect to y1.  The surface is invariant under cyclic relabelling of the
// three systems, so the derivative for system j is d1 of the cyclically
// rotated triple (x_j, x_j+1, x_j+2).  c = c0..c4, a1, a3 over (mu dslip b).
template <typename R>
__device__ __forceinline__ void core_extended(R y1, R y2, R y3, const R *c, R &f, R &d1)
{
    const R arg[12] = {y1 - y2, y2 - y3, y3 - y1,                             // A1..A3
                       2 * y1 - y2 - y3, 2 * y2 - y3 - y1, 2 * y3 - y1 - y2,  // B1..B3
                       3 * y1 - y2 - 2 * y3, 3 * y1 - 2 * y2 - y3, 3 * y2 - y3 - 2 * y1,
                       3 * y2 - 2 * y3 - y1, 3 * y3 - y1 - 2 * y2, 3 * y3 - 2 * y1 - y2}; // D1..D6
    R sn[12], cs[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        if constexpr (sizeof(R) == 8) sincospi(2.0 * (double)arg[q], (double *)&sn[q], (double *)&cs[q]);
        else sincospif(2.0f * (float)arg[q], (float *)&sn[q], (float *)&cs[q]);
    }
    R s2[3], c2[3]; // angles 4 pi A
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        s2[q] = (R)2 * sn[q] * cs[q];
        c2[q] = cs[q] * cs[q] - sn[q] * sn[q];
    }
    const R c0 = c[0], c1 = c[1], cc2 = c[2], c3 = c[3], c4 = c[4], a1 = c[5], a3 = c[6];
    f = c0 + c1 * (cs[0] + cs[1] + cs[2]) + cc2 * (cs[3] + cs[4] + cs[5]) + c3 * (c2[0] + c2[1] + c2[2]) +
        c4 * (cs[6] + cs[7] + cs[8] + cs[9] + cs[10] + cs[11]) + a1 * (sn[0] + sn[1] + sn[2]) +
        a3 * (s2[0] + s2[1] + s2[2]);
    d1 = (R)(2.0 * M_PI) *
         (c1 * (sn[2] - sn[0]) + cc2 * ((R)-2 * sn[3] + sn[4] + sn[5]) + (R)2 * c3 * (s2[2] - s2[0]) +
          c4 * ((R)-3 * sn[6] - (R)3 * sn[7] + (R)2 * sn[8] + sn[9] + sn[10] + (R)2 * sn[11]) +
          a1 * (cs[0] - cs[2]) + (R)2 * a3 * (c2[0] - c2[2]));
}

// Reverse mode of core_extended.  The plane density is a sum of 21 terms
// c_g T(2 pi u.x) over the plane's triple x = (x_0, x_1, x_2), T = cos for
// the groups c1..c4 and sin for a1, a3, with the integer vectors u below (the
// arguments A, B, D of core_extended, and 2A for c3 and a3).  For the
// system k of the triple and mu = the triple's adjoint (masked) values:
//   hv       = sum_i mu_i d2f/dx_k dx_i = -(2 pi)^2 sum_t c_g u_tk (u_t.mu) T_t
//   gc[g]    = mu_k d(df/dx_k)/dc_g     = mu_k sum_{t in g} 2 pi u_tk T'_t
// with T'_t = -sin for cosine terms and cos for sine terms (gc[0], the
// constant c0, stays 0).
template <typename R>
__device__ __forceinline__ void core_extended_adj(const R x[3], const R mu[3], int k, const R *c, R &hv,
                                                  double gc[7])
{
    // u vectors, coefficient index (1..6 = c1..c4, a1, a3) and type (0 cos, 1 sin)
    constexpr signed char U[21][3] = {{1, -1, 0},  {0, 1, -1},  {-1, 0, 1},  {2, -1, -1}, {-1, 2, -1}, {-1, -1, 2},
                                      {2, -2, 0},  {0, 2, -2},  {-2, 0, 2},  {3, -1, -2}, {3, -2, -1}, {-2, 3, -1},
                                      {-1, 3, -2}, {-1, -2, 3}, {-2, -1, 3}, {1, -1, 0},  {0, 1, -1},  {-1, 0, 1},
                                      {2, -2, 0},  {0, 2, -2},  {-2, 0, 2}};
    constexpr signed char G[21] = {1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 4, 4, 4, 5, 5, 5, 6, 6, 6};
    constexpr double tp = 2.0 * M_PI;
    double h = 0.0;
#pragma unroll
    for (int q = 0; q < 7; ++q) gc[q] = 0.0;
#pragma unroll
    for (int t = 0; t < 21; ++t) {
        const R arg = (R)U[t][0] * x[0] + (R)U[t][1] * x[1] + (R)U[t][2] * x[2];
        const R um = (R)U[t][0] * mu[0] + (R)U[t][1] * mu[1] + (R)U[t][2] * mu[2];
        R sn, cs;
        if constexpr (sizeof(R) == 8) sincospi(2.0 * (double)arg, (double *)&sn, (double *)&cs);
        else sincospif(2.0f * (float)arg, (float *)&sn, (float *)&cs);
        const bool is_sin = G[t] >= 5;
        const double Tv = is_sin ? (double)sn : (double)cs;
        const double Td = is_sin ? (double)cs : -(double)sn;
        const double uk = (double)U[t][k];
        h -= tp * tp * (double)c[G[t]] * uk * (double)um * Tv;
        gc[G[t]] += (double)mu[k] * tp * uk * Td;
    }
    hv = (R)h;
}

struct BParams {
    double sig[MAXS * 9]; // sigma_a = C : eps_a
    double eps[MAXS * 9]; // eps_a = b_a n_a^T / dslip
    double four[9];
    double inv2C44;       // 1 / (2 C44)
    double inv1mnu;       // 1 / (1 - xnu)
    double invmu;         // 1 / mu (Voigt mu of Bmatrix)
    int nx, ny, nz, S;
};

// Constants of the on-the-fly Green's-function contraction of the x pass:
// per system the symmetric stress sigma_a = C:eps_a (xx,yy,zz,xy,xz,yz) and
// eps_a folded for a symmetric contraction (xx,yy,zz, xy+yx, xz+zx, yz+zy).
struct GParams {
    double sig6[MAXS * 6];
    double eps6[MAXS * 6];
    double four[9];
    double inv2C44, inv1mnu, invmu;
    double C44x2, C12; // isotropic stiffness: C:Q = 2 C44 Q + C12 tr(Q) I
    double inv_n[3];   // 1/nx, 1/ny, 1/nz (frequencies by multiplication in the x pass)
    int nx, ny, nz;
};

__device__ __forceinline__ double frac_freq(int kk, int n)
{
    // frec(), fft_fftw_slab.cpp:294-322
    return kk >= n / 2 ? (double)(kk - n) / (double)n : (double)kk / (double)n;
}

// ---------------------------------------------------------------- B(k) build
// Packed symmetric B in the point order of T2 (index (k*ny + j)*nx + i), pair
// (a<=b) at slot a*S - a(a-1)/2 + (b-a).  Closed form of Bmatrix's A-tensor
// contraction (fft_fftw_slab.cpp:577-634):
//   mu B_ab = sig_a : eps_b - [2 (u_a . u_b)/f^2 - (u_a.f)(u_b.f)/((1-nu) f^4)] / (2 C44)
// with u_a = sig_a f; B(0) = 0.
template <typename RB>
__global__ void __launch_bounds__(NTHR) k_bbuild(RB *__restrict__ B, BParams P, long long ntot)
{
    const long long q = (long long)blockIdx.x * NTHR + threadIdx.x;
    if (q >= ntot) return;
    const int i = (int)(q % P.nx);
    const long long col = q / P.nx;
    const int j = (int)(col % P.ny);
    const int k = (int)(col / P.ny);
    const double kx = frac_freq(i, P.nx), ky = frac_freq(j, P.ny), kz = frac_freq(k, P.nz);
    double f[3];
    for (int c = 0; c < 3; ++c) f[c] = kx * P.four[0 + c] + ky * P.four[3 + c] + kz * P.four[6 + c];
    const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    const int S = P.S;
    double u[MAXS][3], uf[MAXS];
    for (int a = 0; a < S; ++a) {
        for (int r = 0; r < 3; ++r)
            u[a][r] = P.sig[a * 9 + r * 3 + 0] * f[0] + P.sig[a * 9 + r * 3 + 1] * f[1] +
                      P.sig[a * 9 + r * 3 + 2] * f[2];
        uf[a] = u[a][0] * f[0] + u[a][1] * f[1] + u[a][2] * f[2];
    }
    const double if2 = f2 > 0 ? 1.0 / f2 : 0.0;
    const double if4 = if2 * if2;
    int slot = 0;
    for (int a = 0; a < S; ++a)
        for (int b = a; b < S; ++b, ++slot) {
            double v = 0.0;
            if (f2 > 0) {
                double ce = 0.0;
                for (int m = 0; m < 9; ++m) ce += P.sig[a * 9 + m] * P.eps[b * 9 + m];
                const double uu = u[a][0] * u[b][0] + u[a][1] * u[b][1] + u[a][2] * u[b][2];
                v = (ce - (2.0 * uu * if2 - P.inv1mnu * uf[a] * uf[b] * if4) * P.inv2C44) * P.invmu;
            }
            B[(long long)slot * ntot + q] = (RB)v;
        }
}

// ------------------------------------------------- generic FFT pass (1, 2, 4)
template <typename R, int N, bool INV, int TQ>
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
    fft_stages<N, 1, NL, NTHR, INV>(lds, tw, tid);
    tile_store<N, NL, NTHR>(out, lds, bout, mo.SP, tid);
}

// Load the S lines of NP consecutive columns (T2[a][c0+q][0..N)) into the
// tile, lines ordered (a, q); column-contiguous runs of N elements.
template <int N, int NL, int NP, int NTHR, typename V>
__device__ __forceinline__ void tile_load_cols(const Tile<V> lds, const V *__restrict__ T2, long long c0, long long ncol,
                                               long long ntot, int tid)
{
    constexpr int LDP = NL + 1, TOT = NL * N, PER = (TOT + NTHR - 1) / NTHR;
    if constexpr (sizeof(V) >= 16) {
#pragma unroll 4
        for (int e = tid; e < TOT; e += NTHR) {
            const int l = e / N, p = e % N;
            const int a = l / NP, q = l % NP;
            if (c0 + q < ncol) lds.st(p * LDP + l, ld_stream(&T2[(long long)a * ntot + (c0 + q) * N + p]));
        }
    } else {
        V r[PER];
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int e = tid + t * NTHR;
            const int l = e / N, p = e % N;
            const int a = l / NP, q = l % NP;
            if ((TOT % NTHR == 0 || e < TOT) && c0 + q < ncol) r[t] = ld_stream(&T2[(long long)a * ntot + (c0 + q) * N + p]);
        }
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int e = tid + t * NTHR;
            if (TOT % NTHR == 0 || e < TOT) lds.st((e % N) * LDP + e / N, r[t]);
        }
    }
}

// --------------------------------------------- pass 3: x FFT + contraction
// Columns per x-pass tile: the largest power of two whose S*NP lines of N
// elements stay within `target` elements.  FP64 with 12 systems takes twice
// the tile of the other FP64 cases: at 128^3 x 12 two columns per block (one
// thread per k-point, no duplicated contraction work, 3 blocks per CU) ran
// the pass 12 % faster than one column (two threads per k-point, 5 blocks).
template <typename R, int N, int S> constexpr int mid_pairs()
{
    constexpr int target = sizeof(R) == 8 ? (S == 12 ? 4096 : 2048) : 4096;
    int np = 1;
    while (np * 2 * S * N <= target) np *= 2;
    // the padded tile stays within 60 KB of LDS (the constants and the
    // reduction scratch of the pass come on top)
    while (np > 1 && (long long)N * (S * np + 1) * 2 * (int)sizeof(R) > 61440) np /= 2;
    return np;
}

template <typename R, typename RB, int N, int S>
__global__ void __launch_bounds__(NTHR)
    k_mid(typename cvec<R>::type *__restrict__ T2, const RB *__restrict__ B, long long ncol,
          long long ntot, int bpm, const typename cvec<R>::type *__restrict__ tw, double *__restrict__ part_el,
          const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NP = mid_pairs<R, N, S>();
    constexpr int NL = S * NP, LDP = NL + 1, TOT = NL * N;
    constexpr int NPAIR = S * (S + 1) / 2;
    PFDD_TILE(V, lds, N * LDP);
    __shared__ double red[NTHR / 64];
    const int tid = threadIdx.x;
    // member of a batch (bpm blocks per member), first column of the block
    const int mem = blockIdx.x / bpm;
    const long long c0 = (long long)(blockIdx.x - mem * bpm) * NP;
    T2 += (long long)mem * S * ntot;
    // load: lines (a, q) -> T2[a][c0+q][0..N)
#pragma unroll 4
    for (int e = tid; e < TOT; e += NTHR) {
        const int l = e / N, p = e % N;
        const int a = l / NP, q = l % NP;
        if (c0 + q < ncol) lds.st(p * LDP + l, ld_stream(&T2[(long long)a * ntot + (c0 + q) * N + p]));
    }
    __syncthreads();
    fft_stages<N, 1, NL, NTHR, false>(lds, tw, tid);
    double el = 0.0;
    for (int e = tid; e < NP * N; e += NTHR) {
        const int q = e / N, p = e % N;
        if (c0 + q >= ncol) continue;
        const long long kp = (c0 + q) * N + p;
        V x[S];
#pragma unroll
        for (int b = 0; b < S; ++b) x[b] = lds.ld(p * LDP + b * NP + q);
        R g[S][2];
#pragma unroll
        for (int a = 0; a < S; ++a) g[a][0] = g[a][1] = (R)0;
        int slot = 0;
#pragma unroll
        for (int a = 0; a < S; ++a) {
#pragma unroll
            for (int b = a; b < S; ++b, ++slot) {
                const R bab = (R)B[(long long)slot * ntot + kp];
                g[a][0] += bab * x[b].x;
                g[a][1] += bab * x[b].y;
                if (b != a) {
                    g[b][0] += bab * x[a].x;
                    g[b][1] += bab * x[a].y;
                }
            }
        }
        (void)NPAIR;
#pragma unroll
        for (int a = 0; a < S; ++a) {
            el += (double)x[a].x * (double)g[a][0] + (double)x[a].y * (double)g[a][1];
            lds.st(p * LDP + a * NP + q, mk<V>(g[a][0], g[a][1]));
        }
    }
    __syncthreads();
    fft_stages<N, 1, NL, NTHR, true>(lds, tw, tid);
#pragma unroll 4
    for (int e = tid; e < TOT; e += NTHR) {
        const int l = e / N, p = e % N;
        const int a = l / NP, q = l % NP;
        if (c0 + q < ncol) st_stream(&T2[(long long)a * ntot + (c0 + q) * N + p], lds.ld(p * LDP + l));
    }
    double v[1] = {el};
    block_sum<1, NTHR>(v, red, tid);
    if (tid == 0) part_el[blockIdx.x] = v[0];
}


// x pass with B(k) evaluated on the fly in FP64 instead of streamed from HBM.
// Bmatrix's kernel (fft_fftw_slab.cpp:577-634) is mu B_ab = eps_a:A:eps_b with
// A = C - (C f) G (f C), G = (2 I/f^2 - f f/((1-nu) f^4))/(2 C44).  Applied as
// internal_energy does (:1670-1695) to all systems at once:
//   Sig = sum_b xihat_b sigma_b        (sigma_b = C:eps_b)
//   w = Sig f,  V = f.w
//   Q = [2 sym(w f)/f^2 - V f f/((1-nu) f^4)] / (2 C44)
//   P = Sig - C:Q = Sig - 2 C44 Q - C12 tr(Q) I
//   ghat_a = eps_a : P / mu
// (sigma_a:Q = eps_a:(C:Q) for the isotropic C), and ghat = 0 at f = 0
// (B(0) = 0).  TPP threads share one k-point (wave-uniform split of the
// systems), each producing S/TPP of the outputs.
template <typename R, int N, int S> constexpr int otf_tpp()
{
    constexpr int pts = mid_pairs<R, N, S>() * N;
    return (2 * pts == NTHR && S % 2 == 0) ? 2 : 1;
}

// Scalars of the contraction in the arithmetic type T of the x pass.
template <typename T> struct GConst {
    T inv2C44, inv1mnu, invmu, C44x2, C12;
};

// Hands ghat_a for a = H*SA .. H*SA+SA-1 of one k-point to sink(aa, re, im),
// computed in T (double in the FP64 path, float in the FP32 mode).
template <typename T, int S, int SA, int H, int NP, int LDP, typename V, typename Sink>
__device__ __forceinline__ void otf_point(const Tile<V> lds, int p, int q, const T f[3], T f2,
                                          const T *__restrict__ cst, const GConst<T> &G, Sink sink)
{
    // compiler-only barriers keep the (LDS-broadcast) constants from being
    // hoisted and held in registers across points and systems
    __asm__ volatile("" ::: "memory");
    T Pr[6] = {}, Pi[6] = {}; // Sig, then P
#pragma unroll
    for (int b = 0; b < S; ++b) {
        const V v = lds.ld(p * LDP + b * NP + q);
        const T xr = (T)v.x, xi = (T)v.y;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            Pr[c] += cst[b * 12 + c] * xr;
            Pi[c] += cst[b * 12 + c] * xi;
        }
    }
    if (f2 > 0) {
        // symmetric components xx, yy, zz, xy, xz, yz
        const T wr[3] = {Pr[0] * f[0] + Pr[3] * f[1] + Pr[4] * f[2], Pr[3] * f[0] + Pr[1] * f[1] + Pr[5] * f[2],
                              Pr[4] * f[0] + Pr[5] * f[1] + Pr[2] * f[2]};
        const T wi[3] = {Pi[0] * f[0] + Pi[3] * f[1] + Pi[4] * f[2], Pi[3] * f[0] + Pi[1] * f[1] + Pi[5] * f[2],
                              Pi[4] * f[0] + Pi[5] * f[1] + Pi[2] * f[2]};
        const T Vr = f[0] * wr[0] + f[1] * wr[1] + f[2] * wr[2];
        const T Vi = f[0] * wi[0] + f[1] * wi[1] + f[2] * wi[2];
        const T if2 = (T)1 / f2;
        const T c2 = if2 * G.inv2C44;                 // multiplies (w_i f_j + w_j f_i)
        const T c4 = G.inv1mnu * if2 * if2 * G.inv2C44; // multiplies V f_i f_j
        const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
        T Qr[6], Qi[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const T ff = f[ci[c]] * f[cj[c]];
            Qr[c] = c2 * (wr[ci[c]] * f[cj[c]] + wr[cj[c]] * f[ci[c]]) - c4 * Vr * ff;
            Qi[c] = c2 * (wi[ci[c]] * f[cj[c]] + wi[cj[c]] * f[ci[c]]) - c4 * Vi * ff;
        }
        const T tr_r = G.C12 * (Qr[0] + Qr[1] + Qr[2]), tr_i = G.C12 * (Qi[0] + Qi[1] + Qi[2]);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            Pr[c] -= G.C44x2 * Qr[c] + (c < 3 ? tr_r : (T)0);
            Pi[c] -= G.C44x2 * Qi[c] + (c < 3 ? tr_i : (T)0);
        }
    }
#pragma unroll
    for (int aa = 0; aa < SA; ++aa) {
        __asm__ volatile("" ::: "memory");
        const int a = H * SA + aa;
        const T *ep = cst + a * 12 + 6;
        T gr = 0, gi = 0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            gr += Pr[c] * ep[c];
            gi += Pi[c] * ep[c];
        }
        gr *= G.invmu;
        gi *= G.invmu;
        if (!(f2 > 0)) gr = gi = (T)0;
        sink(aa, gr, gi);
    }
}

// f = four^T k~ of k-point (p, col) (frec, fft_fftw_slab.cpp:280-367), formed
// in FP64 and rounded once to T.
template <int N, int NP, int S, typename T>
__device__ __forceinline__ void otf_freq(const GParams &G, long long col, int p, T f[3], T &f2)
{
    const int c = (int)col; // column (k, j) of one member: ny * nz < 2^31
    const int k = c / G.ny, j = c - k * G.ny;
    // frec's k/n with the reciprocal of n (one rounding more than the
    // division for sizes that are not powers of two): the x pass evaluates it
    // for every k-point of every step
    auto fr = [](int kk, int n, double inv) { return (double)(kk >= n / 2 ? kk - n : kk) * inv; };
    const double kx = fr(p, G.nx, G.inv_n[0]), ky = fr(j, G.ny, G.inv_n[1]), kz = fr(k, G.nz, G.inv_n[2]);
    double fd[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) fd[c] = kx * G.four[c] + ky * G.four[3 + c] + kz * G.four[6 + c];
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = (T)fd[c];
    f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
}

// Minimum waves per SIMD asked of the x pass: 5 for the 12-system FP64 cases
// (their LDS tile allows 5 blocks per CU, and the pass is latency-bound; the
// register budget of 96 then holds without spills), no constraint otherwise.
template <typename R, int N, int S> constexpr int otf_waves()
{
    return (S == 12 && sizeof(R) == 8 && N >= 32 && N <= 128) ? 5 : 1;
}

// The transform of one x-pass tile held in LDS (NP columns x S systems):
// forward FFT along x, on-the-fly contraction ghat_a = sum_b B_ab xihat_b,
// inverse FFT.  Returns this thread's share of sum Re(conj(xihat_a) ghat_a).
template <typename R, int N, int S>
__device__ __forceinline__ double otf_tile(const Tile<typename cvec<R>::type> lds,
                                           const typename cvec<R>::type *__restrict__ tw, const GParams &G,
                                           const GConst<R> &gc, const R *cst, long long c0, long long ncol, int tid)
{
    typedef typename cvec<R>::type V;
    typedef R T;
    constexpr int NP = mid_pairs<R, N, S>();
    constexpr int NL = S * NP, LDP = NL + 1;
    constexpr int TPP = otf_tpp<R, N, S>();
    constexpr int SA = S / TPP;
    constexpr int NPT = NP * N; // k-points per tile
    fft_stages_x<N, 1, NL, NTHR, false>(lds, tw, tid);
    // the elastic-energy term Re(conj(xihat_a) ghat_a) is taken as each ghat_a
    // replaces xihat_a in the tile
    double el = 0.0;
    if constexpr (TPP == 1) {
        // one thread per k-point: it reads all xihat_b of the point, then
        // overwrites them with ghat_a
        for (int pt = tid; pt < NPT; pt += NTHR) {
            const int q = pt / N, p = pt % N;
            if (c0 + q >= ncol) continue;
            T f[3], f2;
            otf_freq<N, NP, S>(G, c0 + q, p, f, f2);
            otf_point<T, S, SA, 0, NP, LDP>(lds, p, q, f, f2, cst, gc, [&](int a, T gr, T gi) {
                const int idx = p * LDP + a * NP + q;
                const V x = lds.ld(idx);
                el += (double)x.x * gr + (double)x.y * gi;
                lds.st(idx, mk<V>((R)gr, (R)gi));
            });
        }
    } else {
        // two waves per set of k-points, each producing half of the systems;
        // all reads of the tile finish (barrier) before any thread writes
        const int h = tid / NPT, pt = tid % NPT;
        const int q = pt / N, p = pt % N;
        const bool ok = c0 + q < ncol;
        T f[3], f2;
        otf_freq<N, NP, S>(G, c0 + q, p, f, f2);
        T g[SA][2];
        auto keep = [&](int aa, T gr, T gi) {
            g[aa][0] = gr;
            g[aa][1] = gi;
        };
        if (ok) {
            if (__builtin_amdgcn_readfirstlane(h) == 0)
                otf_point<T, S, SA, 0, NP, LDP>(lds, p, q, f, f2, cst, gc, keep);
            else
                otf_point<T, S, SA, 1, NP, LDP>(lds, p, q, f, f2, cst, gc, keep);
        }
        __syncthreads();
        if (ok) {
#pragma unroll
            for (int aa = 0; aa < SA; ++aa) {
                const int idx = p * LDP + (h * SA + aa) * NP + q;
                const V x = lds.ld(idx);
                el += (double)x.x * g[aa][0] + (double)x.y * g[aa][1];
                lds.st(idx, mk<V>((R)g[aa][0], (R)g[aa][1]));
            }
        }
    }
    __syncthreads();
    fft_stages_x<N, 1, NL, NTHR, true>(lds, tw, tid);
    return el;
}

template <typename R, int S>
__device__ __forceinline__ void otf_consts(R *cst, GConst<R> &gc, const GParams &G, int tid)
{
    for (int c = tid; c < S * 12; c += NTHR) cst[c] = (R)((c % 12) < 6 ? G.sig6[(c / 12) * 6 + c % 12]
                                                                       : G.eps6[(c / 12) * 6 + c % 12 - 6]);
    gc = GConst<R>{(R)G.inv2C44, (R)G.inv1mnu, (R)G.invmu, (R)G.C44x2, (R)G.C12};
}

// x pass, one tile per block.
template <typename R, int N, int S>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(otf_waves<R, N, S>(), 8)))
    k_mid_otf(typename cvec<R>::type *__restrict__ T2, GParams G, long long ncol, long long ntot, int bpm,
              const typename cvec<R>::type *__restrict__ tw, double *__restrict__ part_el,
              const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NP = mid_pairs<R, N, S>();
    constexpr int NL = S * NP, LDP = NL + 1, TOT = NL * N;
    PFDD_TILE(V, lds, N * LDP);
    __shared__ R cst[S * 12];
    __shared__ double red[NTHR / 64];
    const int tid = threadIdx.x;
    // member of a batch (bpm blocks per member), first column of the block
    const int mem = blockIdx.x / bpm;
    const long long c0 = (long long)(blockIdx.x - mem * bpm) * NP;
    T2 += (long long)mem * S * ntot;
    GConst<R> gc;
    otf_consts<R, S>(cst, gc, G, tid);
    tile_load_cols<N, NL, NP, NTHR>(lds, T2, c0, ncol, ntot, tid);
    __syncthreads();
    double v[1] = {otf_tile<R, N, S>(lds, tw, G, gc, cst, c0, ncol, tid)};
#pragma unroll 4
    for (int e = tid; e < TOT; e += NTHR) {
        const int l = e / N, p = e % N;
        const int a = l / NP, q = l % NP;
        if (c0 + q < ncol) st_stream(&T2[(long long)a * ntot + (c0 + q) * N + p], lds.ld(p * LDP + l));
    }
    block_sum<1, NTHR>(v, red, tid);
    if (tid == 0) part_el[blockIdx.x] = v[0];
}

// x pass, persistent and software-pipelined: each block walks the tiles
// t = blockIdx.x, + gridDim.x, ... and keeps the next tile's input in
// registers, loaded while the current tile is transformed, so that memory
// traffic and the FFT/contraction arithmetic of the resident blocks overlap
// instead of running in lock-step phases.  Tile t is member t / bpm, columns
// (t % bpm) * NP ...; part_el gets one partial per block.
template <int N, int NL, int NP, int NTHR, typename V>
__device__ __forceinline__ void cols_fetch(V *r, const V *__restrict__ T2, long long c0, long long ncol,
                                           long long ntot, int tid)
{
    constexpr int TOT = NL * N, PER = (TOT + NTHR - 1) / NTHR;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
        const int e = tid + t * NTHR;
        const int l = e / N, p = e % N;
        const int a = l / NP, q = l % NP;
        if ((TOT % NTHR == 0 || e < TOT) && c0 + q < ncol) r[t] = ld_stream(&T2[(long long)a * ntot + (c0 + q) * N + p]);
    }
}

template <typename R, int N, int S>
__global__ void __launch_bounds__(NTHR)
    k_mid_otf_pipe(typename cvec<R>::type *__restrict__ T2, GParams G, long long ncol, long long ntot, int bpm,
                   int ntiles, const typename cvec<R>::type *__restrict__ tw, double *__restrict__ part_el,
                   const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NP = mid_pairs<R, N, S>();
    constexpr int NL = S * NP, LDP = NL + 1, TOT = NL * N, PER = (TOT + NTHR - 1) / NTHR;
    PFDD_TILE(V, lds, N * LDP);
    __shared__ R cst[S * 12];
    __shared__ double red[NTHR / 64];
    GConst<R> gc;
    otf_consts<R, S>(cst, gc, G, threadIdx.x);
    int t = blockIdx.x;
    V r[PER];
    {
        const int mem = t / bpm;
        cols_fetch<N, NL, NP, NTHR>(r, T2 + (long long)mem * S * ntot, (long long)(t - mem * bpm) * NP, ncol, ntot,
                                    threadIdx.x);
    }
    double el = 0.0;
    for (; t < ntiles; t += gridDim.x) {
        // the thread index is re-read per tile so that per-element addresses
        // are recomputed rather than held in registers across the loop
        int tid = threadIdx.x;
        __asm__ volatile("" : "+v"(tid));
        const int mem = t / bpm;
        const long long c0 = (long long)(t - mem * bpm) * NP;
        V *T2m = T2 + (long long)mem * S * ntot;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = tid + k * NTHR;
            if (TOT % NTHR == 0 || e < TOT) lds.st((e % N) * LDP + e / N, r[k]);
        }
        __syncthreads();
        const int tn = t + gridDim.x;
        if (tn < ntiles) {
            const int mn = tn / bpm;
            cols_fetch<N, NL, NP, NTHR>(r, T2 + (long long)mn * S * ntot, (long long)(tn - mn * bpm) * NP, ncol,
                                        ntot, tid);
        }
        el += otf_tile<R, N, S>(lds, tw, G, gc, cst, c0, ncol, tid);
#pragma unroll 4
        for (int e = tid; e < TOT; e += NTHR) {
            const int l = e / N, p = e % N;
            const int a = l / NP, q = l % NP;
            if (c0 + q < ncol) T2m[(long long)a * ntot + (c0 + q) * N + p] = lds.ld(p * LDP + l);
        }
        __syncthreads();
    }
    double v[1] = {el};
    block_sum<1, NTHR>(v, red, threadIdx.x);
    if (threadIdx.x == 0) part_el[blockIdx.x] = v[0];
}

// ------------------------------------- pass 5: inverse z FFT + fused update
// update_order_parameter (fft_fftw_slab.cpp:1746-1790) with the core-energy
// derivative of core_energy_* evaluated on the old field, as the reference
// does (it reads data_fftw = xi before the update).  Masked cells keep their
// value; their xinorm term uses the previous unmasked cell's old value in
// (a, i, j, k) order (prev[] precomputed on the host, -1 = the initial 0).
template <typename R, int N, int TQ, bool GEN>
__global__ void __launch_bounds__(NTHR)
    k_update(const typename cvec<R>::type *__restrict__ T1, const typename cvec<R>::type *__restrict__ xo,
             typename cvec<R>::type *__restrict__ xn, const uint8_t *__restrict__ mask,
             const long long *__restrict__ prev, const R *__restrict__ amp, LineMap mi, LineMap mo,
             long long nlines, long long lines_per_sys, StepParams P,
             const typename cvec<R>::type *__restrict__ tw, double *__restrict__ partial,
             long long pstride, const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NL = tile_lines_q<R, N, TQ>(), LDP = NL + 1, TOT = N * NL;
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
        V x;
        x = xold[t];
        // core energy density and its derivative at Re xi of the old field
        R dens = (R)0, dE = (R)0;
        if (GEN && P.model == PFDD_CORE_EXTENDED) {
            // the other two systems of the plane at the same cell
            const int j = lsub[l];
            const long long b0 = idx - (long long)j * P.ncell;
            R y[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) y[q] = q == j ? x.x : xo[b0 + q * P.ncell].x;
            R cc[7];
#pragma unroll
            for (int q = 0; q < 7; ++q) cc[q] = lcoef[q * NL + l];
            R f;
            core_extended<R>(y[j], y[(j + 1) % 3], y[(j + 2) % 3], cc, f, dE);
            dens = f / (R)3; // each of the three systems carries a third of the plane's density
        } else {
            R sp, cp;
            if constexpr (sizeof(R) == 8) sincospi((double)x.x, (double *)&sp, (double *)&cp);
            else sincospif((float)x.x, (float *)&sp, (float *)&cp);
            if (GEN && (P.model == PFDD_CORE_PYRII || P.model == PFDD_CORE_FOURIER)) {
                // a0 + sum_m a_m cos(2 pi m x) + b_m sin(2 pi m x), m <= 4
                const R s1 = (R)2 * sp * cp, c1 = cp * cp - sp * sp;
                R sm = s1, cm = c1;
                dens = (R)P.a0;
#pragma unroll
                for (int m = 1; m <= 4; ++m) {
                    const R am = lcoef[(m - 1) * NL + l], bm = lcoef[(m + 3) * NL + l];
                    dens += am * cm + bm * sm;
                    dE += (R)(2.0 * M_PI * m) * (bm * cm - am * sm);
                    const R sn = sm * c1 + cm * s1, cn = cm * c1 - sm * s1;
                    sm = sn;
                    cm = cn;
                }
                // core_energy_pyrII as executed: its dE_core assignment follows
                // the cell loop (fft_fftw_slab.cpp:1491), so only each system's
                // last cell gets the force
                if (P.model == PFDD_CORE_PYRII && (idx % P.ncell) != P.ncell - 1) dE = (R)0;
            } else {
                // sine series sum_m c_m sin^2(m pi x) (flags 1, 3, 7, 8)
                R sm = sp, cm = cp;
                for (int m = 1; m <= nm; ++m) {
                    const R c = lcoef[(m - 1) * NL + l];
                    dens += c * sm * sm;
                    dE += c * ((R)m * (R)M_PI) * ((R)2 * sm * cm);
                    const R sn = sm * cp + cm * sp, cn = cm * cp - sm * sp;
                    sm = sn;
                    cm = cn;
                }
            }
            R A = (R)1;
            if (amp) A = (R)amp[idx];
            dens *= A;
            dE *= A;
        }
        const R tau = ltau[l];
        V y;
        double dn;
        if (mask && mask[idx]) {
            y = x;
            V pv = mk<V>((R)0, (R)0);
            const long long pi = prev[idx];
            if (pi >= 0) pv = xo[pi];
            const double dr = (double)x.x - (double)pv.x, di = (double)x.y - (double)pv.y;
            dn = dr * dr + di * di;
        } else {
            y.x = x.x - cddt * (g.x * invN - tau + dE);
            y.y = x.y - cddt * (g.y * invN);
            const double dr = (double)y.x - (double)x.x, di = (double)y.y - (double)x.y;
            dn = dr * dr + di * di;
        }
        st_stream(&xn[idx], y);
        s_norm += dn;
        s_ave += (double)y.x;
        s_core += (double)dens;
        s_str += (double)tau * (double)x.x;
    }
    double v[4] = {s_norm, s_ave, s_core, s_str};
    block_sum<4, NTHR>(v, red, tid);
    if (tid == 0) {
        partial[0 * pstride + blockIdx.x] = v[0];
        partial[1 * pstride + blockIdx.x] = v[1];
        partial[2 * pstride + blockIdx.x] = v[2];
        partial[3 * pstride + blockIdx.x] = v[3];
    }
}

// ------------------------------------- character angle (core flag 5)
// gradient() (fft_fftw_slab.cpp:869-1065) and the angle factor of
// core_energy_bcc_perfect (:1071-1125) for one system (blockIdx.y) of the
// field xo: periodic central differences of Re xi, transformed to the global
// axes in the reference's in-place order (the y and z components reuse the
// already transformed x (and y) component, quirk Q6), tang = grad x n,
// theta = acos(tang . b), flipped to pi - theta when (b x tang) . n < 0, and
// "no line" (factor 1) where |tang| <= 0.1.  Writes the factor cof(theta)
// that scales usf' in the update.
struct ThetaParams {
    double four[9];
    double xn[MAXS * 3], xb[MAXS * 3];
    int nx, ny, nz, S;
};

template <typename R, bool THETA = false>
__global__ void __launch_bounds__(NTHR)
    k_theta_cof(const typename cvec<R>::type *__restrict__ xo, R *__restrict__ cof, ThetaParams T, long long N,
                const int *__restrict__ done)
{
    if (*done) return;
    const long long r = (long long)blockIdx.x * NTHR + threadIdx.x;
    if (r >= N) return;
    const int a = blockIdx.y, as = a % T.S;
    const int k = (int)(r % T.nz), j = (int)((r / T.nz) % T.ny), i = (int)(r / ((long long)T.nz * T.ny));
    const typename cvec<R>::type *x = xo + (long long)a * N;
    auto re = [&](int ii, int jj, int kk) {
        return (double)x[((long long)ii * T.ny + jj) * T.nz + kk].x;
    };
    double gx = (re(i + 1 == T.nx ? 0 : i + 1, j, k) - re(i == 0 ? T.nx - 1 : i - 1, j, k)) / 2;
    double gy = (re(i, j + 1 == T.ny ? 0 : j + 1, k) - re(i, j == 0 ? T.ny - 1 : j - 1, k)) / 2;
    double gz = (re(i, j, k + 1 == T.nz ? 0 : k + 1) - re(i, j, k == 0 ? T.nz - 1 : k - 1)) / 2;
    gx = gx * T.four[0] + gy * T.four[3] + gz * T.four[6];
    gy = gx * T.four[1] + gy * T.four[4] + gz * T.four[7];
    gz = gx * T.four[2] + gy * T.four[5] + gz * T.four[8];
    const double *n = T.xn + 3 * as, *b = T.xb + 3 * as;
    double t0 = gy * n[2] - gz * n[1], t1 = gz * n[0] - gx * n[2], t2 = gx * n[1] - gy * n[0];
    const double mag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    double c = 1.0, th = -1.0; // theta = -1: no line through the cell
    if (mag > 0.1) {
        t0 /= mag;
        t1 /= mag;
        t2 /= mag;
        th = acos(t0 * b[0] + t1 * b[1] + t2 * b[2]);
        const double dir = (b[1] * t2 - b[2] * t1) * n[0] + (b[2] * t0 - b[0] * t2) * n[1] + (b[0] * t1 - b[1] * t0) * n[2];
        if (dir < 0.0) th = acos(-1.0) - th;
        c = th < 1.23 ? 1.1603 * th * th - 2.0431 * th + 1 : 0.5473 * th * th - 2.0035 * th + 1.8923;
    }
    cof[(long long)a * N + r] = (R)(THETA ? th : c);
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
    constexpr int NL = tile_lines_q<R, N, TQ>(), LDP = NL + 1, TOT = N * NL;
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
        if constexpr (sizeof(R) == 8) sincospi((double)x.x, (double *)&sp, (double *)&cp);
        else sincospif((float)x.x, (float *)&sp, (float *)&cp);
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

// Adds the per-block adjoint sums of one iteration into gacc[a][slot] in a
// fixed order: one workgroup per system, blocks [a*nbps, (a+1)*nbps).
__global__ void __launch_bounds__(NTHR)
    k_grad_reduce(const double *__restrict__ gpart, long long gstride, int nbps, double *__restrict__ gacc,
                  const int *__restrict__ done)
{
    if (*done) return;
    constexpr int NG = 1 + MAXM;
    __shared__ double red[NG * (NTHR / 64)];
    const int tid = threadIdx.x, a = blockIdx.x;
    double v[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) {
        double sum = 0.0, c = 0.0;
        for (int b = tid; b < nbps; b += NTHR) {
            const double y = gpart[q * gstride + (long long)a * nbps + b] - c;
            const double t = sum + y;
            c = (t - sum) - y;
            sum = t;
        }
        v[q] = sum;
    }
    block_sum<NG, NTHR>(v, red, tid);
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NG; ++q) gacc[a * NG + q] += v[q];
    }
}

// Deterministic reduction of the block partials in two levels: every block
// sums a fixed chunk (Kahan within a thread, fixed-shape tree across the
// block) into stage2; the last block to finish (atomic ticket) sums stage2 in
// fixed order, writes the iteration's scalars, counts the iteration and
// raises `done` on convergence or a non-finite result.
constexpr int NB2 = 32;

__global__ void __launch_bounds__(NTHR)
    k_reduce(const double *__restrict__ partial, long long pstride, int nb5, int nb3, double invN,
             double inv_nx, double tol, double *__restrict__ stage2, unsigned *__restrict__ ticket,
             double *__restrict__ res, int *__restrict__ done, int *__restrict__ steps)
{
    if (*done) return;
    __shared__ double red[NSLOT * (NTHR / 64)];
    __shared__ int last;
    const int tid = threadIdx.x;
    double v[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const int nb = s == 4 ? nb3 : nb5;
        const int chunk = (nb + NB2 - 1) / NB2;
        const int lo = blockIdx.x * chunk, hi = min(nb, lo + chunk);
        double sum = 0.0, c = 0.0;
        for (int b = lo + tid; b < hi; b += NTHR) {
            const double y = partial[s * pstride + b] - c;
            const double t = sum + y;
            c = (t - sum) - y;
            sum = t;
        }
        v[s] = sum;
    }
    block_sum<NSLOT, NTHR>(v, red, tid);
    if (tid == 0) {
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) stage2[s * NB2 + blockIdx.x] = v[s];
        __threadfence();
        last = atomicAdd(ticket, 1u) == NB2 - 1;
    }
    __syncthreads();
    if (!last || tid != 0) return;
    __threadfence();
    double w[NSLOT];
    for (int s = 0; s < NSLOT; ++s) {
        double sum = 0.0, c = 0.0;
        for (int b = 0; b < NB2; ++b) {
            const double y = __builtin_nontemporal_load(&stage2[s * NB2 + b]) - c;
            const double t = sum + y;
            c = (t - sum) - y;
            sum = t;
        }
        w[s] = sum;
    }
    *ticket = 0u;
    const double xinorm = sqrt(w[0]);
    const double e_el = 0.5 * w[4] * invN;
    res[0] = xinorm;
    res[1] = w[1];
    res[2] = e_el;
    res[3] = w[2];
    res[4] = w[3];
    res[5] = e_el + w[2] - w[3];
    res[6] = w[2] * inv_nx;
    const bool finite = isfinite(xinorm) && isfinite(w[1]) && isfinite(e_el) && isfinite(w[2]);
    res[7] = finite ? 0.0 : 1.0;
    *steps += 1;
    if (!finite || (tol > 0 && xinorm < tol)) *done = finite ? 1 : 2;
}

__global__ void k_reset(int *done, int *steps)
{
    if (threadIdx.x == 0) {
        *done = 0;
        *steps = 0;
    }
}

// Per-system sums of Re xi: grid (nb, S); one partial per block.
template <typename R>
__global__ void __launch_bounds__(NTHR)
    k_sys_sums(const typename cvec<R>::type *__restrict__ xi, long long ntot, double *__restrict__ part)
{
    __shared__ double red[NTHR / 64];
    const int a = blockIdx.y;
    const long long chunk = (ntot + gridDim.x - 1) / gridDim.x;
    const long long lo = (long long)blockIdx.x * chunk;
    const long long hi = lo + chunk < ntot ? lo + chunk : ntot;
    double s = 0.0;
    for (long long r = lo + threadIdx.x; r < hi; r += NTHR) s += (double)xi[(long long)a * ntot + r].x;
    double v[1] = {s};
    block_sum<1, NTHR>(v, red, threadIdx.x);
    if (threadIdx.x == 0) part[(long long)a * gridDim.x + blockIdx.x] = v[0];
}

template <typename R>
__global__ void k_from_f64(const double2 *__restrict__ src, typename cvec<R>::type *__restrict__ dst, long long n)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = mk<typename cvec<R>::type>((R)src[q].x, (R)src[q].y);
}
template <typename R>
__global__ void k_to_f64(const typename cvec<R>::type *__restrict__ src, double2 *__restrict__ dst, long long n)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = make_double2((double)src[q].x, (double)src[q].y);
}

// ------------------------------------------------ output-step strain/stress
// strain() (fft_fftw_slab.cpp:3370-3469) builds strain_ij(k) = sum_a F_a,ij(k)
// xihat_a(k) with F_a,ij = f_i (G sigma_a f)_j (Fmatrix, :644-732; G with the
// Voigt mu and nu), i.e. strain_ij = f_i v_j with v = G (Sig f), Sig =
// sum_a xihat_a sigma_a.  Only the symmetric part enters data_eps and
// stress(), so six components (xx,yy,zz,xy,xz,yz) are transformed back.
struct DParams {
    double sig6[MAXS * 6];
    double epsym6[MAXS * 6]; // sym(eps_a): xx,yy,zz,(xy+yx)/2,(xz+zx)/2,(yz+zy)/2
    double four[9];
    double inv2mu, inv1mnu;
    double C44, C12, invN;
    double a_sym[6];    // sym(avepst) - sym(avepsts): the constant of stress()
    double avepst6[6];  // sym(avepst), added to the strain
    double sigmu[9];    // sigma_rot * mu
    int nx, ny, nz, S;
};

template <typename R>
__global__ void __launch_bounds__(NTHR)
    k_diag_comp(const typename cvec<R>::type *__restrict__ xh, typename cvec<R>::type *__restrict__ comp,
                DParams P, long long ntot)
{
    typedef typename cvec<R>::type V;
    const long long q = (long long)blockIdx.x * NTHR + threadIdx.x;
    if (q >= ntot) return;
    const int i = (int)(q % P.nx);
    const long long col = q / P.nx;
    const int j = (int)(col % P.ny), k = (int)(col / P.ny);
    const double kx = frac_freq(i, P.nx), ky = frac_freq(j, P.ny), kz = frac_freq(k, P.nz);
    double f[3];
    for (int c = 0; c < 3; ++c) f[c] = kx * P.four[c] + ky * P.four[3 + c] + kz * P.four[6 + c];
    const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    double Sr[6] = {0, 0, 0, 0, 0, 0}, Si[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < P.S; ++a) {
        const V x = xh[(long long)a * ntot + q];
        for (int c = 0; c < 6; ++c) {
            Sr[c] += P.sig6[a * 6 + c] * (double)x.x;
            Si[c] += P.sig6[a * 6 + c] * (double)x.y;
        }
    }
    const double wr[3] = {Sr[0] * f[0] + Sr[3] * f[1] + Sr[4] * f[2], Sr[3] * f[0] + Sr[1] * f[1] + Sr[5] * f[2],
                          Sr[4] * f[0] + Sr[5] * f[1] + Sr[2] * f[2]};
    const double wi[3] = {Si[0] * f[0] + Si[3] * f[1] + Si[4] * f[2], Si[3] * f[0] + Si[1] * f[1] + Si[5] * f[2],
                          Si[4] * f[0] + Si[5] * f[1] + Si[2] * f[2]};
    const double Vr = f[0] * wr[0] + f[1] * wr[1] + f[2] * wr[2];
    const double Vi = f[0] * wi[0] + f[1] * wi[1] + f[2] * wi[2];
    double vr[3] = {0, 0, 0}, vi[3] = {0, 0, 0};
    if (f2 > 0) {
        const double if2 = 1.0 / f2, c4 = P.inv1mnu * if2 * if2;
        for (int r = 0; r < 3; ++r) {
            vr[r] = (2.0 * wr[r] * if2 - f[r] * Vr * c4) * P.inv2mu;
            vi[r] = (2.0 * wi[r] * if2 - f[r] * Vi * c4) * P.inv2mu;
        }
    }
    const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 6; ++c) {
        const double er = 0.5 * (f[ci[c]] * vr[cj[c]] + f[cj[c]] * vr[ci[c]]);
        const double ei = 0.5 * (f[ci[c]] * vi[cj[c]] + f[cj[c]] * vi[ci[c]]);
        comp[(long long)c * ntot + q] = mk<V>((R)er, (R)ei);
    }
}

// Real space: data_eps (symmetric strain + sym(avepst)) and data_sigma =
// C:(strain + avepst - eps_d(xi_post) - avepsts) + sigma_rot mu (stress(),
// fft_fftw_slab.cpp:3474-3553), written [component (i + 3j)][cell].
template <typename R>
__global__ void __launch_bounds__(NTHR)
    k_diag_real(const typename cvec<R>::type *__restrict__ comp, const typename cvec<R>::type *__restrict__ xi,
                double2 *__restrict__ eps, double *__restrict__ sig, DParams P, long long ntot)
{
    const long long r = (long long)blockIdx.x * NTHR + threadIdx.x;
    if (r >= ntot) return;
    double er[6], ei[6], x[6];
    for (int c = 0; c < 6; ++c) {
        const auto v = comp[(long long)c * ntot + r];
        er[c] = (double)v.x * P.invN;
        ei[c] = (double)v.y * P.invN;
        x[c] = er[c] + P.a_sym[c];
    }
    for (int a = 0; a < P.S; ++a) {
        const double xr = (double)xi[(long long)a * ntot + r].x;
        for (int c = 0; c < 6; ++c) x[c] -= P.epsym6[a * 6 + c] * xr;
    }
    const double tr = x[0] + x[1] + x[2];
    const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 6; ++c) {
        const int ii = ci[c], jj = cj[c];
        const double2 e = make_double2(er[c] + P.avepst6[c], ei[c]);
        const double sv = 2.0 * P.C44 * x[c] + (ii == jj ? P.C12 * tr : 0.0);
        eps[(long long)(ii + 3 * jj) * ntot + r] = e;
        eps[(long long)(jj + 3 * ii) * ntot + r] = e;
        sig[(long long)(ii + 3 * jj) * ntot + r] = sv + P.sigmu[3 * ii + jj];
        sig[(long long)(jj + 3 * ii) * ntot + r] = sv + P.sigmu[3 * jj + ii];
    }
}

// ------------------------------------------------------------------ dispatch
#define PFDD_SIZES(X) X(8) X(16) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(512)

bool size_supported(int n)
{
#define PFDD_CASE(n_) if (n == n_) return true;
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return false;
}

bool systems_supported(int s) { return s == 1 || s == 2 || s == 3 || s == 4 || s == 6 || s == 12; }

template <typename R>
int tile_lines_rt(int n, int tq)
{
#define PFDD_CASE(n_)                                              \
    if (n == n_) {                                                 \
        if (tq == 0) return tile_lines_q<R, n_, 0>();              \
        if (tq == 2) return tile_lines_q<R, n_, 2>();              \
        return tile_lines_q<R, n_, 1>();                           \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return 0;
}

// Dispatch of a tile-shape choice tq (0, 1, 2) to the template argument.
#define PFDD_TQ_DISPATCH(tq, ...)                                  \
    switch (tq) {                                                  \
    case 0: { constexpr int TQ = 0; __VA_ARGS__; } break;          \
    case 2: { constexpr int TQ = 2; __VA_ARGS__; } break;          \
    default: { constexpr int TQ = 1; __VA_ARGS__; } break;         \
    }

template <typename R, bool INV>
hipError_t launch_pass(int n, const void *in, void *out, LineMap mi, LineMap mo, long long nlines,
                       const void *tw, const int *done, hipStream_t st, const uint8_t *zmask = nullptr,
                       int tq = 1)
{
    typedef typename cvec<R>::type V;
#define PFDD_CASE(n_)                                                                               \
    if (n == n_) {                                                                                  \
        PFDD_TQ_DISPATCH(tq, {                                                                      \
            constexpr int NL = tile_lines_q<R, n_, TQ>();                                           \
            const unsigned nb = (unsigned)((nlines + NL - 1) / NL);                                 \
            PFDD_LAUNCH((k_fft_pass<R, n_, INV, TQ>), dim3(nb), dim3(NTHR), st,                     \
                               (const V *)in, (V *)out, mi, mo, nlines, (const V *)tw, done, zmask); \
        })                                                                                          \
        return hipGetLastError();                                                                   \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
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
            constexpr int NL = tile_lines_q<R, n_, 0>();                                             \
            const unsigned nb = (unsigned)((nlines + NL - 1) / NL);                                  \
            PFDD_LAUNCH((k_update<R, n_, 0, true>), dim3(nb), dim3(NTHR), st, (const V *)T1,         \
                        (const V *)xo, (V *)xn, mask, prev, (const R *)amp, mi, mo, nlines, lps, P,  \
                        (const V *)tw, partial, pstride, done);                                      \
            return hipGetLastError();                                                                \
        }                                                                                            \
        PFDD_TQ_DISPATCH(tq, {                                                                       \
            constexpr int NL = tile_lines_q<R, n_, TQ>();                                            \
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
            constexpr int NL = tile_lines_q<R, n_, TQ>();                                              \
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

template <typename R, int S>
int mid_blocks_rt(int n, long long ncol)
{
#define PFDD_CASE(n_) if (n == n_) return (int)((ncol + mid_pairs<R, n_, S>() - 1) / mid_pairs<R, n_, S>());
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return 0;
}

template <typename R>
int mid_blocks(int n, int S, long long ncol)
{
    switch (S) {
    case 1: return mid_blocks_rt<R, 1>(n, ncol);
    case 2: return mid_blocks_rt<R, 2>(n, ncol);
    case 3: return mid_blocks_rt<R, 3>(n, ncol);
    case 4: return mid_blocks_rt<R, 4>(n, ncol);
    case 6: return mid_blocks_rt<R, 6>(n, ncol);
    case 12: return mid_blocks_rt<R, 12>(n, ncol);
    }
    return 0;
}

template <typename R, int N, int S> constexpr bool mid_fits()
{
    return (long long)N * (S * mid_pairs<R, N, S>() + 1) * 2 * sizeof(R) <= 61440;
}

template <typename R, int S>
hipError_t launch_mid_s(int n, int M, void *T2, const void *B, long long ncol, long long ntot, const void *tw,
                        double *part_el, const int *done, hipStream_t st)
{
    typedef typename cvec<R>::type V;
#define PFDD_CASE(n_)                                                                             \
    if (n == n_) {                                                                                \
        if constexpr (mid_fits<R, n_, S>()) {                                                     \
            constexpr int NP = mid_pairs<R, n_, S>();                                             \
            const int bpm = (int)((ncol + NP - 1) / NP);                                          \
            PFDD_LAUNCH((k_mid<R, R, n_, S>), dim3((unsigned)(M * bpm)), dim3(NTHR), st,          \
                               (V *)T2, (const R *)B, ncol, ntot, bpm, (const V *)tw, part_el, done); \
            return hipGetLastError();                                                             \
        } else {                                                                                  \
            return hipErrorInvalidValue;                                                          \
        }                                                                                         \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

template <typename R>
bool mid_supported(int n, int S)
{
    auto chk = [&](auto sconst) -> bool {
        constexpr int SS = decltype(sconst)::value;
#define PFDD_CASE(n_) if (n == n_) return mid_fits<R, n_, SS>();
        PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
        return false;
    };
    switch (S) {
    case 1: return chk(std::integral_constant<int, 1>());
    case 2: return chk(std::integral_constant<int, 2>());
    case 3: return chk(std::integral_constant<int, 3>());
    case 4: return chk(std::integral_constant<int, 4>());
    case 6: return chk(std::integral_constant<int, 6>());
    case 12: return chk(std::integral_constant<int, 12>());
    }
    return false;
}

template <typename R>
hipError_t launch_mid(int n, int S, int M, void *T2, const void *B, long long ncol, long long ntot, const void *tw,
                      double *part_el, const int *done, hipStream_t st)
{
    switch (S) {
    case 1: return launch_mid_s<R, 1>(n, M, T2, B, ncol, ntot, tw, part_el, done, st);
    case 2: return launch_mid_s<R, 2>(n, M, T2, B, ncol, ntot, tw, part_el, done, st);
    case 3: return launch_mid_s<R, 3>(n, M, T2, B, ncol, ntot, tw, part_el, done, st);
    case 4: return launch_mid_s<R, 4>(n, M, T2, B, ncol, ntot, tw, part_el, done, st);
    case 6: return launch_mid_s<R, 6>(n, M, T2, B, ncol, ntot, tw, part_el, done, st);
    case 12: return launch_mid_s<R, 12>(n, M, T2, B, ncol, ntot, tw, part_el, done, st);
    }
    return hipErrorInvalidValue;
}

template <typename R, int S>
hipError_t launch_mid_otf_s(int n, int M, void *T2, const GParams &G, long long ncol, long long ntot,
                            const void *tw, double *part_el, const int *done, hipStream_t st, int pipe, int ncu,
                            int *nblocks)
{
    typedef typename cvec<R>::type V;
#define PFDD_CASE(n_)                                                                             \
    if (n == n_) {                                                                                \
        if constexpr (mid_fits<R, n_, S>()) {                                                     \
            constexpr int NP = mid_pairs<R, n_, S>();                                             \
            const int bpm = (int)((ncol + NP - 1) / NP);                                          \
            if (!pipe) {                                                                          \
                *nblocks = M * bpm;                                                               \
                PFDD_LAUNCH((k_mid_otf<R, n_, S>), dim3((unsigned)(M * bpm)), dim3(NTHR), st,      \
                            (V *)T2, G, ncol, ntot, bpm, (const V *)tw, part_el, done);           \
            } else {                                                                              \
                /* as many blocks as fit on the device at once */                                 \
                static const int occ = [] {                                                       \
                    int o = 0;                                                                    \
                    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                             \
                            &o, (const void *)k_mid_otf_pipe<R, n_, S>, NTHR, 0) != hipSuccess)   \
                        o = 1;                                                                    \
                    return o > 0 ? o : 1;                                                         \
                }();                                                                              \
                const long long nt = (long long)M * bpm, cap = (long long)ncu * occ;             \
                *nblocks = (int)(nt < cap ? nt : cap);                                            \
                PFDD_LAUNCH((k_mid_otf_pipe<R, n_, S>), dim3((unsigned)*nblocks), dim3(NTHR), st,  \
                            (V *)T2, G, ncol, ntot, bpm, (int)nt, (const V *)tw, part_el, done);  \
            }                                                                                     \
            return hipGetLastError();                                                             \
        } else {                                                                                  \
            return hipErrorInvalidValue;                                                          \
        }                                                                                         \
    }
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_mid_otf(int n, int S, int M, void *T2, const GParams &G, long long ncol, long long ntot,
                          const void *tw, double *part_el, const int *done, hipStream_t st, int pipe, int ncu,
                          int *nblocks)
{
    switch (S) {
    case 1: return launch_mid_otf_s<R, 1>(n, M, T2, G, ncol, ntot, tw, part_el, done, st, pipe, ncu, nblocks);
    case 2: return launch_mid_otf_s<R, 2>(n, M, T2, G, ncol, ntot, tw, part_el, done, st, pipe, ncu, nblocks);
    case 3: return launch_mid_otf_s<R, 3>(n, M, T2, G, ncol, ntot, tw, part_el, done, st, pipe, ncu, nblocks);
    case 4: return launch_mid_otf_s<R, 4>(n, M, T2, G, ncol, ntot, tw, part_el, done, st, pipe, ncu, nblocks);
    case 6: return launch_mid_otf_s<R, 6>(n, M, T2, G, ncol, ntot, tw, part_el, done, st, pipe, ncu, nblocks);
    case 12: return launch_mid_otf_s<R, 12>(n, M, T2, G, ncol, ntot, tw, part_el, done, st, pipe, ncu, nblocks);
    }
    return hipErrorInvalidValue;
}

} // namespace

// ================================================================== handle
struct pfdd_gl {
    int nx = 0, ny = 0, nz = 0, S = 0, prec = PFDD_FP64;
    int M = 1, SM = 0;      // members of a batch, systems over all members (S*M)
    double *taud = nullptr; // device tau[SM] of a batch
    void *cof = nullptr;    // core flag 5: per-element angle factor (R), refreshed every iteration
    unsigned version = 0;   // bumped when kernel arguments change (tau, core coefficients)
    long long N = 0;
    size_t esz = 16; // bytes per complex element
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    void *xi[2] = {nullptr, nullptr};
    int cur = 0;
    void *T1 = nullptr;
    void *B = nullptr;
    void *tw[3] = {nullptr, nullptr, nullptr}; // x, y, z twiddle tables
    uint8_t *mask = nullptr;
    long long *prev = nullptr;
    void *amp = nullptr;
    double *partial = nullptr;
    long long pstride = 0;
    int nb3 = 0, nb5 = 0, nb5max = 0; // blocks of pass 3 / pass 5 (current tile, smallest tile)
    int tq[4] = {1, 1, 1, 1};          // tile shape of fwd z, fwd y, inv y, inv z + update
    int xpipe = 0;                     // x pass: 0 one tile per block, 1 persistent pipelined
    ThetaParams TP{};                  // core flag 5
    int ncu = 0;                       // compute units of the device
    double *stage2 = nullptr; // NSLOT * NB2
    unsigned *ticket = nullptr;
    bool stored_B = false;    // stream packed B(k) from HBM instead of evaluating it on the fly
    GParams G{};
    double *res = nullptr; // NRES doubles
    int *done = nullptr;
    int *steps = nullptr;
    StepParams P{};
    // host copies for diagnostics
    double xn[MAXS * 3], xbv[MAXS * 3], eps[MAXS * 9];
    double C44 = 0, C12 = 0, C11 = 0, S12 = 0, S44 = 0, dslip = 0, mu = 0, xnu = 0;
    double four[9];
    bool stepped = false; // a step ran since the last set_xi (xi[1-cur] holds the pre-update field)
    void *dw[2] = {nullptr, nullptr}; // output-step field workspace, max(S,6)*N complex each
    double2 *deps = nullptr;          // 9*N strain field (complex)
    double *dsig = nullptr;           // 9*N stress field (real)
    // reverse mode: checkpoints of a recorded run and the adjoint workspace
    struct Tape {
        int nsteps = 0, every = 0;
        bool valid = false;
        StepParams P{};             // tau and core coefficients the run used
        std::vector<void *> ckpt;   // xi at steps 0, every, 2*every, ...
        std::vector<void *> seg;    // recomputed states of one segment (every - 1)
        void *lam[2] = {nullptr, nullptr};
        double *gpart = nullptr;    // (1 + MAXM) * nb5max per-block sums
        double *gacc = nullptr;     // S * (1 + MAXM) accumulated sums
    } tape;
};

namespace {

void make_twiddles(int n, std::vector<double> &tw)
{
    tw.resize(2 * n);
    for (int m = 0; m < n; ++m) {
        // exp(-2 pi i m / n), evaluated with the angle folded to the first octant
        long long mm = m;
        double c, s;
        const long long n8 = 8LL * mm;
        // use sincospi-like exactness at multiples of n/8
        if (n8 % n == 0) {
            const int oct = (int)(n8 / n);
            static const double r2 = 0.70710678118654752440;
            const double cs[8] = {1, r2, 0, -r2, -1, -r2, 0, r2};
            const double sn[8] = {0, r2, 1, r2, 0, -r2, -1, -r2};
            c = cs[oct];
            s = sn[oct];
        } else {
            const double ang = 2.0 * M_PI * (double)mm / (double)n;
            c = std::cos(ang);
            s = std::sin(ang);
        }
        tw[2 * m] = c;
        tw[2 * m + 1] = -s;
    }
}

int upload_twiddles(pfdd_gl *h)
{
    const int dims[3] = {h->nx, h->ny, h->nz};
    for (int d = 0; d < 3; ++d) {
        std::vector<double> tw;
        make_twiddles(dims[d], tw);
        const size_t bytes = (size_t)dims[d] * h->esz;
        HIPCHK(hipMalloc(&h->tw[d], bytes));
        if (h->prec == PFDD_FP64) {
            HIPCHK(hipMemcpy(h->tw[d], tw.data(), bytes, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(tw.begin(), tw.end());
            HIPCHK(hipMemcpy(h->tw[d], f.data(), bytes, hipMemcpyHostToDevice));
        }
    }
    return PFDD_OK;
}

// Line maps of the passes (see the file header).
struct Maps {
    LineMap p1i, p1o, p2i, p2o, p5i, p5o;
    long long n1, n2, n5;
};

Maps make_maps_dims(long long nx, long long ny, long long nz, long long S)
{
    Maps m;
    // pass 1: lines (a,i | j), FFT along k
    m.p1i = {ny * nz, 0, nz, 1, ny, 1};
    m.p1o = {nz * ny, 0, 1, ny, ny, 1};
    m.n1 = S * nx * ny;
    // pass 2 / 4: lines (a | k | i), FFT along j; T1[a][i][k][j] <-> T2[a][k][j][i]
    m.p2i = {nx * nz * ny, ny, nz * ny, 1, nx, nz};
    m.p2o = {nz * ny * nx, ny * nx, 1, nx, nx, nz};
    m.n2 = S * nz * nx;
    // pass 5: lines (a,i | j), FFT along k; T1[a][i][k][j] -> xi[a][i][j][k]
    m.p5i = {nz * ny, 0, 1, ny, ny, 1};
    m.p5o = {ny * nz, 0, nz, 1, ny, 1};
    m.n5 = S * nx * ny;
    return m;
}

Maps make_maps(const pfdd_gl *h, long long S) { return make_maps_dims(h->nx, h->ny, h->nz, S); }

// One GL iteration from xo into T2 (T2 doubles as the temporary of passes
// 2-4), with the stress and core parameters P.
template <typename R>
int step_core(pfdd_gl *h, double tol, hipEvent_t *ev, const void *xo, void *T2, const StepParams &P)
{
    const Maps m = make_maps(h, h->SM);
    hipStream_t st = h->stream;
    // ev (profiling): 2 events per kernel, [2k] at its start and [2k+1] at its end
    auto arm = [&](int k) {
        if (ev) {
            t_ev_start = ev[2 * k];
            t_ev_stop = ev[2 * k + 1];
        }
    };
    struct Disarm {
        ~Disarm() { t_ev_start = t_ev_stop = nullptr; }
    } disarm;
    if (h->cof) {
        // core flag 5: gradient() + the angle factor on the old field
        hipLaunchKernelGGL((k_theta_cof<R>), dim3((unsigned)((h->N + NTHR - 1) / NTHR), h->SM), dim3(NTHR), 0, st,
                           (const typename cvec<R>::type *)xo, (R *)h->cof, h->TP, h->N, h->done);
        HIPCHK(hipGetLastError());
    }
    arm(0);
    HIPCHK((launch_pass<R, false>(h->nz, xo, h->T1, m.p1i, m.p1o, m.n1, h->tw[2], h->done, st, nullptr,
                                  h->tq[0])));
    arm(1);
    HIPCHK((launch_pass<R, false>(h->ny, h->T1, T2, m.p2i, m.p2o, m.n2, h->tw[1], h->done, st, nullptr,
                                  h->tq[1])));
    arm(2);
    if (h->stored_B)
        HIPCHK((launch_mid<R>(h->nx, h->S, h->M, T2, h->B, (long long)h->ny * h->nz, h->N, h->tw[0],
                              h->partial + 4 * h->pstride, h->done, st)));
    else
        HIPCHK((launch_mid_otf<R>(h->nx, h->S, h->M, T2, h->G, (long long)h->ny * h->nz, h->N, h->tw[0],
                                  h->partial + 4 * h->pstride, h->done, st, h->xpipe, h->ncu, &h->nb3)));
    arm(3);
    HIPCHK((launch_pass<R, true>(h->ny, T2, h->T1, m.p2o, m.p2i, m.n2, h->tw[1], h->done, st, nullptr,
                                 h->tq[2])));
    arm(4);
    HIPCHK(launch_update<R>(h->nz, h->tq[3], h->T1, xo, T2, h->mask, h->prev, h->cof ? h->cof : h->amp, m.p5i,
                            m.p5o, m.n5, (long long)h->nx * h->ny, P, h->tw[2], h->partial, h->pstride, h->done,
                            st));
    arm(5);
    // E_core as the reference accumulates it (sum / nx); core_energy_extended
    // and core_energy_pyrII do not accumulate it (fft_fftw_slab.cpp:1320,1483)
    const bool no_eref = P.model == PFDD_CORE_EXTENDED || P.model == PFDD_CORE_PYRII;
    PFDD_LAUNCH(k_reduce, dim3(NB2), dim3(NTHR), st, h->partial, h->pstride, h->nb5, h->nb3, 1.0 / (double)h->N,
                no_eref ? 0.0 : 1.0 / (double)h->nx, tol, h->stage2, h->ticket, h->res, h->done, h->steps);
    HIPCHK(hipGetLastError());
    return PFDD_OK;
}

template <typename R>
int step_once(pfdd_gl *h, double tol, hipEvent_t *ev)
{
    const int rc = step_core<R>(h, tol, ev, h->xi[h->cur], h->xi[1 - h->cur], h->P);
    if (rc) return rc;
    h->cur = 1 - h->cur;
    h->stepped = true;
    return PFDD_OK;
}

// ---------------------------------------------------------------- autotune
// tq[0..3]: tile shapes of the line passes, tq[4]: x-pass variant
void set_tiles(pfdd_gl *h, const int tq[5])
{
    for (int p = 0; p < 4; ++p) h->tq[p] = tq[p];
    // the update of the coupled and Fourier core models has one tile shape,
    // the smallest (fewest values per thread: no spills in FP32)
    const int m = h->P.model;
    if (m == PFDD_CORE_EXTENDED || m == PFDD_CORE_PYRII || m == PFDD_CORE_FOURIER) h->tq[3] = 0;
    h->xpipe = h->stored_B ? 0 : tq[4];
    const long long n5 = (long long)h->SM * h->nx * h->ny;
    const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, h->tq[3])
                                        : tile_lines_rt<float>(h->nz, h->tq[3]);
    h->nb5 = (int)((n5 + nl - 1) / nl);
    ++h->version; // captured graphs hold the old launch shapes
}

// Per-pass device time (ms) of the five tunable kernels over nsteps GL
// iterations with the handle's current choices.
template <typename R>
int time_passes(pfdd_gl *h, int nsteps, double acc[5])
{
    hipEvent_t ev[12];
    for (int q = 0; q < 12; ++q) HIPCHK(hipEventCreate(&ev[q]));
    for (int p = 0; p < 5; ++p) acc[p] = 0.0;
    const int pass_k[5] = {0, 1, 3, 4, 2}; // fwd z, fwd y, inv y, inv z + update, x pass
    int rc = PFDD_OK;
    for (int s = 0; s < nsteps && !rc; ++s) {
        rc = step_once<R>(h, 0.0, ev);
        if (rc) break;
        if (hipEventSynchronize(ev[11]) != hipSuccess) {
            rc = fail(PFDD_ERR_HIP, "autotune timing failed");
            break;
        }
        for (int p = 0; p < 5; ++p) {
            float t = 0;
            hipEventElapsedTime(&t, ev[2 * pass_k[p]], ev[2 * pass_k[p] + 1]);
            acc[p] += t;
        }
    }
    for (int q = 0; q < 12; ++q) (void)hipEventDestroy(ev[q]);
    return rc;
}

struct TuneKey {
    int nx, ny, nz, S, M, prec, flags;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(nx, ny, nz, S, M, prec, flags) < std::tie(o.nx, o.ny, o.nz, o.S, o.M, o.prec, o.flags);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::array<int, 5>> g_tune_cache;

// Times every tile shape (0, 1, 2) of each line pass and both x-pass variants
// and keeps the fastest per kernel.  The field, ping-pong parity and step
// counters are restored.  Choices that would break the reverse mode's
// one-system-per-block rule for the update pass are skipped.
template <typename R>
int autotune(pfdd_gl *h, int nsteps, int choice[5])
{
    const TuneKey key{h->nx, h->ny, h->nz, h->S, h->M, h->prec,
                      (h->stored_B ? 1 : 0) | (h->mask ? 2 : 0) | (h->amp ? 4 : 0)};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            set_tiles(h, it->second.data());
            if (choice)
                for (int p = 0; p < 5; ++p) choice[p] = it->second[p];
            return PFDD_OK;
        }
    }
    const size_t fb = (size_t)h->SM * h->N * h->esz;
    void *save = nullptr;
    HIPCHK(hipMalloc(&save, fb));
    const int cur0 = h->cur;
    const bool stepped0 = h->stepped;
    HIPCHK(hipMemcpyAsync(save, h->xi[cur0], fb, hipMemcpyDeviceToDevice, h->stream));
    int arg[5] = {1, 1, 1, 1, 0};
    int rc = PFDD_OK;
    // candidates: tile shape q = 0, 1, 2 with the one-tile-per-block x pass,
    // then (on-the-fly contraction only) the pipelined x pass.  The clock of
    // an idle GPU ramps up over the first few dozen dispatches, so all
    // candidates are timed after a common warm-up, in two interleaved rounds,
    // keeping each kernel's best time per candidate.
    const int ncand = h->stored_B ? 3 : 4;
    double tmin[4][5];
    for (int c = 0; c < 4; ++c)
        for (int p = 0; p < 5; ++p) tmin[c][p] = 1e300;
    {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
        double warm[5];
        rc = time_passes<R>(h, 20, warm);
    }
    for (int round = 0; round < 2 && !rc; ++round)
        for (int c = 0; c < ncand && !rc; ++c) {
            const int all[5] = {c < 3 ? c : 1, c < 3 ? c : 1, c < 3 ? c : 1, c < 3 ? c : 1, c < 3 ? 0 : 1};
            set_tiles(h, all);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
            double warm[5], t[5];
            if ((rc = time_passes<R>(h, 1, warm)) || (rc = time_passes<R>(h, nsteps, t))) break;
            for (int p = 0; p < 5; ++p) tmin[c][p] = t[p] < tmin[c][p] ? t[p] : tmin[c][p];
        }
    if (!rc) {
        double best[5] = {1e300, 1e300, 1e300, 1e300, 1e300};
        for (int q = 0; q < 3; ++q) {
            const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, q) : tile_lines_rt<float>(h->nz, q);
            const bool upd_ok = ((long long)h->nx * h->ny) % nl == 0;
            for (int p = 0; p < 4; ++p)
                if (tmin[q][p] < best[p] && (p < 3 || upd_ok)) {
                    best[p] = tmin[q][p];
                    arg[p] = q;
                }
            best[4] = tmin[q][4] < best[4] ? tmin[q][4] : best[4];
        }
        if (ncand == 4 && tmin[3][4] < best[4]) arg[4] = 1;
    }
    HIPCHK(hipMemcpyAsync(h->xi[cur0], save, fb, hipMemcpyDeviceToDevice, h->stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(save));
    h->cur = cur0;
    h->stepped = stepped0;
    if (rc) {
        const int dflt[5] = {1, 1, 1, 1, 0};
        set_tiles(h, dflt);
        return rc;
    }
    set_tiles(h, arg);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {arg[0], arg[1], arg[2], arg[3], arg[4]};
    }
    if (choice)
        for (int p = 0; p < 5; ++p) choice[p] = arg[p];
    return PFDD_OK;
}

// One reverse-mode iteration: lam_in = dL/dxi_{n+1} -> lam_out = dL/dxi_n,
// with xs = xi_n, accumulating dL/dtau and dL/dcoef sums into tape.gacc.
template <typename R>
int adjoint_core(pfdd_gl *h, const void *xs, const void *lam_in, void *lam_out, const StepParams &P)
{
    const Maps m = make_maps(h, h->SM);
    hipStream_t st = h->stream;
    HIPCHK((launch_pass<R, false>(h->nz, lam_in, h->T1, m.p1i, m.p1o, m.n1, h->tw[2], h->done, st, h->mask,
                                  h->tq[0])));
    HIPCHK((launch_pass<R, false>(h->ny, h->T1, lam_out, m.p2i, m.p2o, m.n2, h->tw[1], h->done, st, nullptr,
                                  h->tq[1])));
    if (h->stored_B)
        HIPCHK((launch_mid<R>(h->nx, h->S, h->M, lam_out, h->B, (long long)h->ny * h->nz, h->N, h->tw[0],
                              h->partial + 4 * h->pstride, h->done, st)));
    else
        HIPCHK((launch_mid_otf<R>(h->nx, h->S, h->M, lam_out, h->G, (long long)h->ny * h->nz, h->N, h->tw[0],
                                  h->partial + 4 * h->pstride, h->done, st, h->xpipe, h->ncu, &h->nb3)));
    HIPCHK((launch_pass<R, true>(h->ny, lam_out, h->T1, m.p2o, m.p2i, m.n2, h->tw[1], h->done, st, nullptr,
                                 h->tq[2])));
    const long long gstride = h->nb5max;
    HIPCHK(launch_adj_update<R>(h->nz, h->tq[3], h->T1, lam_in, xs, lam_out, h->mask, h->amp, m.p5i, m.p5o,
                                m.n5, (long long)h->nx * h->ny, P, h->tw[2], h->tape.gpart, gstride, h->done,
                                st));
    hipLaunchKernelGGL(k_grad_reduce, dim3(h->S), dim3(NTHR), 0, st, h->tape.gpart, gstride, h->nb5 / h->S,
                       h->tape.gacc, h->done);
    HIPCHK(hipGetLastError());
    return PFDD_OK;
}

int check_handle(const pfdd_gl *h)
{
    if (!h) return fail(PFDD_ERR_ARG, "null handle");
    return PFDD_OK;
}

int set_device(const pfdd_gl *h)
{
    HIPCHK(hipSetDevice(h->device));
    return PFDD_OK;
}

void destroy_all(pfdd_gl *h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->cof, h->taud, h->xi[0], h->xi[1], h->T1, h->B, h->tw[0], h->tw[1], h->tw[2], h->mask, h->prev,
                    h->amp, h->partial, h->res, h->done, h->steps, h->stage2, h->ticket, h->dw[0], h->dw[1],
                    h->deps, h->dsig, h->tape.lam[0], h->tape.lam[1], h->tape.gpart, h->tape.gacc};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (void *p : h->tape.ckpt)
        if (p) hipFree(p);
    for (void *p : h->tape.seg)
        if (p) hipFree(p);
    if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
    delete h;
}

int set_coef(pfdd_gl *h, int model, int ncoef, const double *coef, bool has_amp)
{
    StepParams &P = h->P;
    for (int q = 0; q < MAXS * MAXM; ++q) P.coef[q] = 0.0;
    P.a0 = 0.0;
    P.model = model;
    switch (model) {
    case PFDD_CORE_NONE:
        P.ncoef = 0;
        break;
    case PFDD_CORE_PERFECT:
        if (!coef) return fail(PFDD_ERR_ARG, "core model 1 needs coef[0] = An'");
        P.ncoef = 1;
        for (int a = 0; a < h->S; ++a) P.coef[a * MAXM] = coef[0];
        break;
    case PFDD_CORE_SINE:
        if (!coef) return fail(PFDD_ERR_ARG, "core model 3 needs coef[0..1] = isf', Cn'");
        P.ncoef = 2;
        for (int a = 0; a < h->S; ++a) {
            P.coef[a * MAXM] = coef[0];
            P.coef[a * MAXM + 1] = coef[1];
        }
        break;
    case PFDD_CORE_MPEA:
        if (!has_amp) return fail(PFDD_ERR_ARG, "core model 7 needs the per-cell amplitude amp[]");
        P.ncoef = 1;
        for (int a = 0; a < h->S; ++a) P.coef[a * MAXM] = 1.0;
        break;
    case PFDD_CORE_BCC:
        if (!coef) return fail(PFDD_ERR_ARG, "core model 5 needs coef[0] = usf'");
        if (has_amp) return fail(PFDD_ERR_UNSUPPORTED, "core model 5 takes no amp[]");
        P.ncoef = 1;
        for (int a = 0; a < h->S; ++a) P.coef[a * MAXM] = coef[0];
        break;
    case PFDD_CORE_EXTENDED:
        if (!coef) return fail(PFDD_ERR_ARG, "core model 2 needs coef[0..6] = c0..c4, a1, a3");
        if (h->S % 3) return fail(PFDD_ERR_UNSUPPORTED, "core model 2 couples systems in threes (ns %% 3 == 0)");
        P.ncoef = 7;
        for (int a = 0; a < h->S; ++a)
            for (int m = 0; m < 7; ++m) P.coef[a * MAXM + m] = coef[m];
        break;
    case PFDD_CORE_PYRII:
    case PFDD_CORE_FOURIER:
        if (!coef) return fail(PFDD_ERR_ARG, "core models 4/9 need coef[0..8] = aa0..aa4, bb1..bb4");
        P.ncoef = 8;
        P.a0 = coef[0];
        for (int a = 0; a < h->S; ++a)
            for (int m = 0; m < 8; ++m) P.coef[a * MAXM + m] = coef[1 + m];
        break;
    case PFDD_CORE_SERIES:
        if (ncoef < 0 || ncoef > MAXM) return fail(PFDD_ERR_ARG, "ncoef must be in [0, %d]", MAXM);
        if (ncoef > 0 && !coef) return fail(PFDD_ERR_ARG, "core model 8 needs coef[ns*ncoef]");
        P.ncoef = ncoef;
        for (int a = 0; a < h->S; ++a)
            for (int m = 0; m < ncoef; ++m) P.coef[a * MAXM + m] = coef[a * ncoef + m];
        break;
    default:
        return fail(PFDD_ERR_ARG, "unknown core model %d", model);
    }
    return PFDD_OK;
}

template <typename R>
int build_B(pfdd_gl *h)
{
    BParams bp{};
    const int S = h->S;
    for (int a = 0; a < S; ++a) {
        double e[9], s[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) e[3 * i + j] = h->xbv[3 * a + i] * h->xn[3 * a + j] / h->dslip;
        const double tr = e[0] + e[4] + e[8];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                s[3 * i + j] = h->C44 * (e[3 * i + j] + e[3 * j + i]) + (i == j ? h->C12 * tr : 0.0);
        for (int q = 0; q < 9; ++q) {
            bp.eps[a * 9 + q] = e[q];
            bp.sig[a * 9 + q] = s[q];
            h->eps[a * 9 + q] = e[q];
        }
    }
    for (int q = 0; q < 9; ++q) bp.four[q] = h->four[q];
    bp.inv2C44 = 1.0 / (2.0 * h->C44);
    bp.inv1mnu = 1.0 / (1.0 - h->xnu);
    bp.invmu = 1.0 / h->mu;
    bp.nx = h->nx;
    bp.ny = h->ny;
    bp.nz = h->nz;
    bp.S = S;
    GParams &G = h->G;
    static const int si[6] = {0, 4, 8, 1, 2, 5}; // xx yy zz xy xz yz in row-major 3x3
    for (int a = 0; a < S; ++a)
        for (int c = 0; c < 6; ++c) {
            const int q = si[c], r = q / 3, cc = q % 3;
            G.sig6[a * 6 + c] = bp.sig[a * 9 + q];
            G.eps6[a * 6 + c] = c < 3 ? bp.eps[a * 9 + q] : bp.eps[a * 9 + 3 * r + cc] + bp.eps[a * 9 + 3 * cc + r];
        }
    for (int q = 0; q < 9; ++q) G.four[q] = h->four[q];
    G.C44x2 = 2.0 * h->C44;
    G.C12 = h->C12;
    G.inv2C44 = bp.inv2C44;
    G.inv1mnu = bp.inv1mnu;
    G.invmu = bp.invmu;
    G.nx = h->nx;
    G.ny = h->ny;
    G.nz = h->nz;
    G.inv_n[0] = 1.0 / h->nx;
    G.inv_n[1] = 1.0 / h->ny;
    G.inv_n[2] = 1.0 / h->nz;
    if (!h->stored_B) return PFDD_OK;
    const long long np = (long long)S * (S + 1) / 2;
    HIPCHK(hipMalloc(&h->B, (size_t)np * h->N * sizeof(R)));
    const unsigned nb = (unsigned)((h->N + NTHR - 1) / NTHR);
    hipLaunchKernelGGL((k_bbuild<R>), dim3(nb), dim3(NTHR), 0, h->stream, (R *)h->B, bp, h->N);
    HIPCHK(hipGetLastError());
    return PFDD_OK;
}

int upload_field_f64(pfdd_gl *h, const double *host, void *dst)
{
    const long long n = (long long)h->SM * h->N;
    if (h->prec == PFDD_FP64) {
        HIPCHK(hipMemcpyAsync(dst, host, (size_t)n * 16, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return PFDD_OK;
    }
    // FP32: stage through T1 as FP64 would not fit (T1 is S*N float2); convert in chunks
    const long long chunk = 1LL << 22;
    void *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, (size_t)chunk * 16));
    for (long long o = 0; o < n; o += chunk) {
        const long long c = (n - o) < chunk ? (n - o) : chunk;
        HIPCHK(hipMemcpyAsync(tmp, host + 2 * o, (size_t)c * 16, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL((k_from_f64<float>), dim3((unsigned)((c + 255) / 256)), dim3(256), 0, h->stream,
                           (const double2 *)tmp, (float2 *)dst + o, c);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(tmp));
    return PFDD_OK;
}

int download_field_f64(pfdd_gl *h, const void *src, double *host)
{
    const long long n = (long long)h->SM * h->N;
    if (h->prec == PFDD_FP64) {
        HIPCHK(hipMemcpyAsync(host, src, (size_t)n * 16, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return PFDD_OK;
    }
    const long long chunk = 1LL << 22;
    void *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, (size_t)chunk * 16));
    for (long long o = 0; o < n; o += chunk) {
        const long long c = (n - o) < chunk ? (n - o) : chunk;
        hipLaunchKernelGGL((k_to_f64<float>), dim3((unsigned)((c + 255) / 256)), dim3(256), 0, h->stream,
                           (const float2 *)src + o, (double2 *)tmp, c);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host + 2 * o, tmp, (size_t)c * 16, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(tmp));
    return PFDD_OK;
}

template <typename R>
int diag_fields(pfdd_gl *h, const double *sigma_rot, const double *avepst, const double *avepsts,
                double *data_eps, double *data_sigma)
{
    const long long S = h->S, N = h->N, SC = S > 6 ? S : 6;
    const size_t wbytes = (size_t)SC * N * h->esz;
    for (int q = 0; q < 2; ++q)
        if (!h->dw[q]) HIPCHK(hipMalloc(&h->dw[q], wbytes));
    if (!h->deps) HIPCHK(hipMalloc(&h->deps, sizeof(double2) * 9 * N));
    if (!h->dsig) HIPCHK(hipMalloc(&h->dsig, sizeof(double) * 9 * N));
    DParams P{};
    const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
    for (int a = 0; a < S; ++a)
        for (int c = 0; c < 6; ++c) {
            P.sig6[a * 6 + c] = h->G.sig6[a * 6 + c];
            P.epsym6[a * 6 + c] = 0.5 * (h->eps[a * 9 + 3 * ci[c] + cj[c]] + h->eps[a * 9 + 3 * cj[c] + ci[c]]);
        }
    for (int q = 0; q < 9; ++q) {
        P.four[q] = h->four[q];
        P.sigmu[q] = sigma_rot[q] * h->mu;
    }
    for (int c = 0; c < 6; ++c) {
        const int i = ci[c], j = cj[c];
        P.avepst6[c] = 0.5 * (avepst[3 * i + j] + avepst[3 * j + i]);
        P.a_sym[c] = P.avepst6[c] - 0.5 * (avepsts[3 * i + j] + avepsts[3 * j + i]);
    }
    P.inv2mu = 1.0 / (2.0 * h->mu);
    P.inv1mnu = 1.0 / (1.0 - h->xnu);
    P.C44 = h->C44;
    P.C12 = h->C12;
    P.invN = 1.0 / (double)N;
    P.nx = h->nx;
    P.ny = h->ny;
    P.nz = h->nz;
    P.S = h->S;
    // strain() reads the k-space field of the pre-update xi (quirk Q2)
    const void *pre = h->stepped ? h->xi[1 - h->cur] : h->xi[h->cur];
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, st, h->done, h->steps);
    HIPCHK(hipGetLastError());
    const Maps ms = make_maps(h, S), m6 = make_maps(h, 6);
    const long long ncol = (long long)h->ny * h->nz;
    const LineMap xs = {h->nx, 0, 0, 1, 1, 1}; // x lines of the [.][k][j][i] layout
    HIPCHK((launch_pass<R, false>(h->nz, pre, h->dw[0], ms.p1i, ms.p1o, ms.n1, h->tw[2], h->done, st)));
    HIPCHK((launch_pass<R, false>(h->ny, h->dw[0], h->dw[1], ms.p2i, ms.p2o, ms.n2, h->tw[1], h->done, st)));
    HIPCHK((launch_pass<R, false>(h->nx, h->dw[1], h->dw[0], xs, xs, S * ncol, h->tw[0], h->done, st)));
    const unsigned nb = (unsigned)((N + NTHR - 1) / NTHR);
    typedef typename cvec<R>::type V;
    hipLaunchKernelGGL((k_diag_comp<R>), dim3(nb), dim3(NTHR), 0, st, (const V *)h->dw[0], (V *)h->dw[1], P, N);
    HIPCHK(hipGetLastError());
    HIPCHK((launch_pass<R, true>(h->nx, h->dw[1], h->dw[0], xs, xs, 6 * ncol, h->tw[0], h->done, st)));
    HIPCHK((launch_pass<R, true>(h->ny, h->dw[0], h->dw[1], m6.p2o, m6.p2i, m6.n2, h->tw[1], h->done, st)));
    HIPCHK((launch_pass<R, true>(h->nz, h->dw[1], h->dw[0], m6.p5i, m6.p5o, m6.n5, h->tw[2], h->done, st)));
    hipLaunchKernelGGL((k_diag_real<R>), dim3(nb), dim3(NTHR), 0, st, (const V *)h->dw[0],
                       (const V *)h->xi[h->cur], h->deps, h->dsig, P, N);
    HIPCHK(hipGetLastError());
    if (data_eps) HIPCHK(hipMemcpyAsync(data_eps, h->deps, sizeof(double2) * 9 * N, hipMemcpyDeviceToHost, st));
    std::vector<double> sg;
    if (data_sigma) {
        sg.resize((size_t)9 * N);
        HIPCHK(hipMemcpyAsync(sg.data(), h->dsig, sizeof(double) * 9 * N, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    if (data_sigma)
        for (long long q = 0; q < 9 * N; ++q) {
            data_sigma[2 * q] = sg[q];
            data_sigma[2 * q + 1] = 0.0;
        }
    return PFDD_OK;
}

int collect_steps(pfdd_gl *h, int cur0, pfdd_gl_reduce *out)
{
    double res[NRES];
    int done = 0, steps = 0;
    HIPCHK(hipMemcpyAsync(res, h->res, sizeof(res), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(&done, h->done, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(&steps, h->steps, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // the field moved steps times from buffer cur0
    h->cur = (steps % 2) ? 1 - cur0 : cur0;
    if (done == 2)
        return fail(PFDD_ERR_NONFINITE, "non-finite reduction after %d GL iterations (field diverged)", steps);
    if (out) {
        out->xinorm = res[0];
        out->xiave = res[1];
        out->e_elastic = res[2];
        out->e_core = res[3];
        out->e_stress = res[4];
        out->e_total = res[5];
        out->e_core_ref = res[6];
        out->steps_done = steps;
        out->converged = done == 1;
    }
    return PFDD_OK;
}

int tape_backward_any(pfdd_gl *h, const void *dxi_T, bool dev_in, void *dxi_0, bool dev_out, double *dtau,
                             double *dcoef)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    pfdd_gl::Tape &T = h->tape;
    if (!T.valid) return fail(PFDD_ERR_STATE, "no recorded run (call pfdd_gl_tape_record first)");
    if (!dxi_T) return fail(PFDD_ERR_ARG, "null dxi_T");
    const size_t fb = (size_t)h->S * h->N * h->esz;
    for (int q = 0; q < 2; ++q)
        if (!T.lam[q]) HIPCHK(hipMalloc(&T.lam[q], fb));
    while ((int)T.seg.size() < T.every - 1) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, fb));
        T.seg.push_back(p);
    }
    constexpr int NG = 1 + MAXM;
    if (!T.gpart) HIPCHK(hipMalloc((void **)&T.gpart, sizeof(double) * NG * h->nb5max));
    if (!T.gacc) HIPCHK(hipMalloc((void **)&T.gacc, sizeof(double) * NG * MAXS));
    const long long n = (long long)h->S * h->N;
    // seed
    if (dev_in) {
        if (h->prec == PFDD_FP64)
            HIPCHK(hipMemcpyAsync(T.lam[0], dxi_T, fb, hipMemcpyDeviceToDevice, h->stream));
        else {
            hipLaunchKernelGGL((k_from_f64<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                               (const double2 *)dxi_T, (float2 *)T.lam[0], n);
            HIPCHK(hipGetLastError());
        }
    } else if ((rc = upload_field_f64(h, (const double *)dxi_T, T.lam[0]))) {
        return rc;
    }
    HIPCHK(hipMemsetAsync(T.gacc, 0, sizeof(double) * NG * MAXS, h->stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
    HIPCHK(hipGetLastError());
    int cur = 0;
    const int nseg = (T.nsteps + T.every - 1) / T.every;
    std::vector<const void *> xs(T.every);
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * T.every, len = std::min(T.every, T.nsteps - s0);
        xs[0] = T.ckpt[sg];
        for (int j = 1; j < len; ++j) {
            rc = h->prec == PFDD_FP64 ? step_core<double>(h, 0.0, nullptr, xs[j - 1], T.seg[j - 1], T.P)
                                      : step_core<float>(h, 0.0, nullptr, xs[j - 1], T.seg[j - 1], T.P);
            if (rc) return rc;
            xs[j] = T.seg[j - 1];
        }
        for (int j = len - 1; j >= 0; --j) {
            rc = h->prec == PFDD_FP64 ? adjoint_core<double>(h, xs[j], T.lam[cur], T.lam[1 - cur], T.P)
                                      : adjoint_core<float>(h, xs[j], T.lam[cur], T.lam[1 - cur], T.P);
            if (rc) return rc;
            cur = 1 - cur;
        }
    }
    double g[NG * MAXS];
    int done = 0;
    HIPCHK(hipMemcpyAsync(g, T.gacc, sizeof(g), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(&done, h->done, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (done) return fail(PFDD_ERR_NONFINITE, "non-finite state while recomputing the recorded run");
    const double c = T.P.cddt;
    for (int a = 0; a < h->S; ++a) {
        if (dtau) dtau[a] = c * g[a * NG];
        for (int m = 1; m <= T.P.ncoef; ++m)
            if (dcoef) dcoef[a * T.P.ncoef + m - 1] = -c * g[a * NG + m];
    }
    for (int q = 0; q < NG * h->S; ++q)
        if (!std::isfinite(g[q])) return fail(PFDD_ERR_NONFINITE, "non-finite adjoint sums");
    if (dxi_0) {
        if (dev_out) {
            if (h->prec == PFDD_FP64)
                HIPCHK(hipMemcpyAsync(dxi_0, T.lam[cur], fb, hipMemcpyDeviceToDevice, h->stream));
            else {
                hipLaunchKernelGGL((k_to_f64<float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                                   (const float2 *)T.lam[cur], (double2 *)dxi_0, n);
                HIPCHK(hipGetLastError());
            }
            HIPCHK(hipStreamSynchronize(h->stream));
        } else if ((rc = download_field_f64(h, T.lam[cur], (double *)dxi_0))) {
            return rc;
        }
    }
    return PFDD_OK;
}

} // namespace

// ===================================================================== C ABI
extern "C" {

int pfdd_gl_version(void) { return (0 << 16) | 1; }

const char *pfdd_gl_last_error(void) { return g_err.c_str(); }

int pfdd_gl_create(const pfdd_gl_desc *d, pfdd_gl **out) { return pfdd_gl_create_batch(d, 1, out); }

int pfdd_gl_create_batch(const pfdd_gl_desc *d, int nbatch, pfdd_gl **out)
{
    g_err.clear();
    if (!d || !out) return fail(PFDD_ERR_ARG, "null descriptor or output");
    *out = nullptr;
    if (nbatch < 1) return fail(PFDD_ERR_ARG, "nbatch must be >= 1");
    if ((long long)nbatch * d->ns > 4096) return fail(PFDD_ERR_ARG, "nbatch * ns must be <= 4096");
    if (!size_supported(d->nx) || !size_supported(d->ny) || !size_supported(d->nz))
        return fail(PFDD_ERR_UNSUPPORTED, "grid %dx%dx%d: each dimension must be one of 8,16,32,40,48,64,80,96,128,160,192,256,512",
                    d->nx, d->ny, d->nz);
    if (!systems_supported(d->ns))
        return fail(PFDD_ERR_UNSUPPORTED, "slip_systems=%d: supported counts are 1,2,3,4,6,12", d->ns);
    if (d->precision != PFDD_FP64 && d->precision != PFDD_FP32)
        return fail(PFDD_ERR_ARG, "precision must be PFDD_FP64 or PFDD_FP32");
    const bool ok_mid = d->precision == PFDD_FP64 ? mid_supported<double>(d->nx, d->ns)
                                                  : mid_supported<float>(d->nx, d->ns);
    if (!ok_mid)
        return fail(PFDD_ERR_UNSUPPORTED, "nx=%d with %d systems exceeds the x-pass LDS tile", d->nx, d->ns);
    if (!d->xn || !d->xb) return fail(PFDD_ERR_ARG, "xn/xb are required");
    if (!(d->C44 > 0) || !(d->dslip != 0)) return fail(PFDD_ERR_ARG, "C44 must be > 0 and dslip != 0");

    pfdd_gl *h = new pfdd_gl();
    h->nx = d->nx;
    h->ny = d->ny;
    h->nz = d->nz;
    h->S = d->ns;
    h->M = nbatch;
    h->SM = nbatch * d->ns;
    h->prec = d->precision;
    h->N = (long long)d->nx * d->ny * d->nz;
    h->esz = d->precision == PFDD_FP64 ? 16 : 8;
    h->device = d->device;
    if (hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, d->device) != hipSuccess ||
        h->ncu <= 0)
        h->ncu = 256;
    for (int q = 0; q < 9; ++q) h->four[q] = d->four[q];
    for (int q = 0; q < 3 * d->ns; ++q) {
        h->xn[q] = d->xn[q];
        h->xbv[q] = d->xb[q];
    }
    h->C44 = d->C44;
    h->C12 = d->C12;
    h->C11 = d->C11;
    h->S12 = d->S12;
    h->S44 = d->S44;
    h->dslip = d->dslip;
    // Voigt moduli as Bmatrix recomputes them (fft_fftw_slab.cpp:487-490)
    h->mu = d->C44 - (2.0 * d->C44 + d->C12 - d->C11) / 5.0;
    const double ll = d->C12 - (2.0 * d->C44 + d->C12 - d->C11) / 5.0;
    const double young = h->mu * (3 * ll + 2 * h->mu) / (ll + h->mu);
    h->xnu = young / 2.0 / h->mu - 1.0;
    h->stored_B = (d->flags & PFDD_GL_FLAG_STORED_B) != 0;
    h->P.S = d->ns;
    h->P.ncell = h->N;
    h->P.cddt = d->cddt;
    h->P.invN = 1.0 / (double)h->N;

    auto bail = [&](int rc) {
        destroy_all(h);
        return rc;
    };
    int rc = set_coef(h, d->core_model, d->ncoef, d->coef, d->amp != nullptr);
    if (rc) return bail(rc);
    hipError_t e = hipSetDevice(d->device);
    if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipSetDevice(%d): %s", d->device, hipGetErrorString(e)));
    if (d->stream) {
        h->stream = (hipStream_t)d->stream;
    } else {
        e = hipStreamCreate(&h->stream);
        if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
        h->own_stream = true;
    }
    const size_t fbytes = (size_t)h->SM * h->N * h->esz;
    auto alloc = [&](void **p, size_t bytes) -> int {
        hipError_t ee = hipMalloc(p, bytes);
        if (ee != hipSuccess)
            return fail(PFDD_ERR_HIP, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(ee));
        return PFDD_OK;
    };
    if ((rc = alloc(&h->xi[0], fbytes)) || (rc = alloc(&h->xi[1], fbytes)) || (rc = alloc(&h->T1, fbytes)))
        return bail(rc);
    e = hipMemsetAsync(h->xi[0], 0, fbytes, h->stream);
    if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipMemset: %s", hipGetErrorString(e)));
    if ((rc = upload_twiddles(h))) return bail(rc);

    // per-block partials
    const long long n5 = (long long)h->SM * h->nx * h->ny;
    const int nl5 = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, 1) : tile_lines_rt<float>(h->nz, 1);
    const int nl5min = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, 0) : tile_lines_rt<float>(h->nz, 0);
    h->nb5 = (int)((n5 + nl5 - 1) / nl5);
    h->nb5max = (int)((n5 + nl5min - 1) / nl5min);
    h->nb3 = h->M * (h->prec == PFDD_FP64 ? mid_blocks<double>(h->nx, h->S, (long long)h->ny * h->nz)
                                          : mid_blocks<float>(h->nx, h->S, (long long)h->ny * h->nz));
    if (h->M > 1) {
        if ((rc = alloc((void **)&h->taud, sizeof(double) * h->SM))) return bail(rc);
        e = hipMemsetAsync(h->taud, 0, sizeof(double) * h->SM, h->stream);
        if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipMemset: %s", hipGetErrorString(e)));
        h->P.taud = h->taud;
    }
    h->pstride = h->nb5max > h->nb3 ? h->nb5max : h->nb3;
    {
        const int dflt[5] = {1, 1, 1, 1, 0};
        set_tiles(h, dflt); // nb5 of the model's update tile
    }
    if ((rc = alloc((void **)&h->partial, sizeof(double) * NSLOT * h->pstride))) return bail(rc);
    if ((rc = alloc((void **)&h->res, sizeof(double) * NRES))) return bail(rc);
    if ((rc = alloc((void **)&h->stage2, sizeof(double) * NSLOT * NB2)) ||
        (rc = alloc((void **)&h->ticket, sizeof(unsigned))))
        return bail(rc);
    e = hipMemsetAsync(h->ticket, 0, sizeof(unsigned), h->stream);
    if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipMemset: %s", hipGetErrorString(e)));
    if ((rc = alloc((void **)&h->done, sizeof(int))) || (rc = alloc((void **)&h->steps, sizeof(int))))
        return bail(rc);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);

    // obstacle mask + quirk Q3 predecessor table (restarting with every member)
    if (d->xo) {
        const long long n = (long long)h->SM * h->N, nmem = (long long)h->S * h->N;
        std::vector<uint8_t> m(n);
        std::vector<long long> pv(n);
        long long last = -1;
        bool any = false;
        for (long long q = 0; q < n; ++q) {
            if (q % nmem == 0) last = -1;
            m[q] = d->xo[q] != 0.0;
            any |= m[q] != 0;
            pv[q] = last;
            if (!m[q]) last = q;
        }
        if (any) {
            if ((rc = alloc((void **)&h->mask, n)) || (rc = alloc((void **)&h->prev, n * sizeof(long long))))
                return bail(rc);
            if (hipMemcpy(h->mask, m.data(), n, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(h->prev, pv.data(), n * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(PFDD_ERR_HIP, "mask upload failed"));
        }
    }
    if (d->amp) {
        const long long n = (long long)h->SM * h->N;
        if (h->prec == PFDD_FP64) {
            if ((rc = alloc(&h->amp, n * sizeof(double)))) return bail(rc);
            if (hipMemcpy(h->amp, d->amp, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(PFDD_ERR_HIP, "amp upload failed"));
        } else {
            std::vector<float> f(d->amp, d->amp + n);
            if ((rc = alloc(&h->amp, n * sizeof(float)))) return bail(rc);
            if (hipMemcpy(h->amp, f.data(), n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(PFDD_ERR_HIP, "amp upload failed"));
        }
    }
    // gradient()'s geometry: the character angle of core flag 5 and of
    // pfdd_gl_theta
    for (int q = 0; q < 9; ++q) h->TP.four[q] = d->four[q];
    for (int q = 0; q < 3 * d->ns; ++q) {
        h->TP.xn[q] = d->xn[q];
        h->TP.xb[q] = d->xb[q];
    }
    h->TP.nx = h->nx;
    h->TP.ny = h->ny;
    h->TP.nz = h->nz;
    h->TP.S = h->S;
    if (d->core_model == PFDD_CORE_BCC)
        if ((rc = alloc(&h->cof, (size_t)h->SM * h->N * (h->prec == PFDD_FP64 ? 8 : 4)))) return bail(rc);
    rc = h->prec == PFDD_FP64 ? build_B<double>(h) : build_B<float>(h);
    if (rc) return bail(rc);
    e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "setup: %s", hipGetErrorString(e)));
    if (d->flags & PFDD_GL_FLAG_AUTOTUNE) {
        rc = h->prec == PFDD_FP64 ? autotune<double>(h, 4, nullptr) : autotune<float>(h, 4, nullptr);
        if (rc) return bail(rc);
    }
    *out = h;
    return PFDD_OK;
}

int pfdd_gl_theta(pfdd_gl *h, double *theta)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!theta) return fail(PFDD_ERR_ARG, "null theta");
    // gradient() runs at the start of an iteration on the field before its
    // update: the previous ping-pong buffer once the handle has stepped
    const void *pre = h->stepped ? h->xi[1 - h->cur] : h->xi[h->cur];
    const size_t n = (size_t)h->SM * h->N;
    void *buf = nullptr;
    HIPCHK(hipMalloc(&buf, n * sizeof(double)));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
    if (h->prec == PFDD_FP64)
        hipLaunchKernelGGL((k_theta_cof<double, true>), dim3((unsigned)((h->N + NTHR - 1) / NTHR), h->SM), dim3(NTHR),
                           0, h->stream, (const double2 *)pre, (double *)buf, h->TP, h->N, h->done);
    else
        hipLaunchKernelGGL((k_theta_cof<float, true>), dim3((unsigned)((h->N + NTHR - 1) / NTHR), h->SM), dim3(NTHR),
                           0, h->stream, (const float2 *)pre, (float *)buf, h->TP, h->N, h->done);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) {
        if (h->prec == PFDD_FP64) {
            e = hipMemcpy(theta, buf, n * sizeof(double), hipMemcpyDeviceToHost);
        } else {
            std::vector<float> f(n);
            e = hipMemcpy(f.data(), buf, n * sizeof(float), hipMemcpyDeviceToHost);
            for (size_t q = 0; q < n; ++q) theta[q] = f[q];
        }
    }
    hipFree(buf);
    if (e != hipSuccess) return fail(PFDD_ERR_HIP, "theta: %s", hipGetErrorString(e));
    return PFDD_OK;
}

int pfdd_gl_set_tuning(pfdd_gl *h, const int choice[5])
{
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!choice) return fail(PFDD_ERR_ARG, "null choice");
    for (int p = 0; p < 4; ++p)
        if (choice[p] < 0 || choice[p] > 2) return fail(PFDD_ERR_ARG, "tile shape choice must be 0, 1 or 2");
    if (choice[4] < 0 || choice[4] > 1) return fail(PFDD_ERR_ARG, "x-pass choice must be 0 or 1");
    const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, choice[3])
                                        : tile_lines_rt<float>(h->nz, choice[3]);
    if (h->tape.valid && ((long long)h->nx * h->ny) % nl != 0)
        return fail(PFDD_ERR_UNSUPPORTED, "update tile shape incompatible with the recorded reverse-mode run");
    set_tiles(h, choice);
    return PFDD_OK;
}

int pfdd_gl_autotune(pfdd_gl *h, int nsteps, int choice[5])
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (nsteps < 1) return fail(PFDD_ERR_ARG, "nsteps must be >= 1");
    return h->prec == PFDD_FP64 ? autotune<double>(h, nsteps, choice) : autotune<float>(h, nsteps, choice);
}

int pfdd_gl_destroy(pfdd_gl *h)
{
    if (!h) return PFDD_OK;
    destroy_all(h);
    return PFDD_OK;
}

int pfdd_gl_set_xi(pfdd_gl *h, const double *xi)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!xi) return fail(PFDD_ERR_ARG, "null xi");
    h->stepped = false;
    return upload_field_f64(h, xi, h->xi[h->cur]);
}

int pfdd_gl_get_xi(pfdd_gl *h, double *xi)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!xi) return fail(PFDD_ERR_ARG, "null xi");
    return download_field_f64(h, h->xi[h->cur], xi);
}

int pfdd_gl_xi_device(pfdd_gl *h, void **dptr)
{
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!dptr) return fail(PFDD_ERR_ARG, "null output");
    *dptr = h->xi[h->cur];
    return PFDD_OK;
}

int pfdd_gl_set_xi_device(pfdd_gl *h, const void *src)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!src) return fail(PFDD_ERR_ARG, "null source");
    HIPCHK(hipMemcpyAsync(h->xi[h->cur], src, (size_t)h->SM * h->N * h->esz, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->stepped = false;
    return PFDD_OK;
}

int pfdd_gl_get_xi_device(pfdd_gl *h, void *dst)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!dst) return fail(PFDD_ERR_ARG, "null destination");
    HIPCHK(hipMemcpyAsync(dst, h->xi[h->cur], (size_t)h->SM * h->N * h->esz, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return PFDD_OK;
}

int pfdd_gl_set_tau(pfdd_gl *h, const double *tau)
{
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!tau) return fail(PFDD_ERR_ARG, "null tau");
    if (h->M > 1) {
        if ((rc = set_device(h))) return rc;
        HIPCHK(hipMemcpyAsync(h->taud, tau, sizeof(double) * h->SM, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return PFDD_OK;
    }
    for (int a = 0; a < h->S; ++a) h->P.tau[a] = tau[a];
    ++h->version;
    return PFDD_OK;
}

int pfdd_gl_set_core_coef(pfdd_gl *h, int ncoef, const double *coef)
{
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (h->cof) return fail(PFDD_ERR_UNSUPPORTED, "core model 5 keeps its usf' (set at create)");
    ++h->version;
    return set_coef(h, PFDD_CORE_SERIES, ncoef, coef, h->amp != nullptr);
}

int pfdd_gl_step(pfdd_gl *h, int nsteps, double tol, pfdd_gl_reduce *out)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (nsteps < 0) return fail(PFDD_ERR_ARG, "nsteps < 0");
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
    HIPCHK(hipGetLastError());
    const int cur0 = h->cur;
    for (int s = 0; s < nsteps; ++s) {
        rc = h->prec == PFDD_FP64 ? step_once<double>(h, tol, nullptr) : step_once<float>(h, tol, nullptr);
        if (rc) return rc;
    }
    if (!out && tol <= 0) return PFDD_OK;
    return collect_steps(h, cur0, out);
}

int pfdd_gl_tape_record(pfdd_gl *h, int nsteps, int every, pfdd_gl_reduce *out)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (nsteps < 1 || every < 1) return fail(PFDD_ERR_ARG, "need nsteps >= 1 and every >= 1");
    if (h->M > 1) return fail(PFDD_ERR_UNSUPPORTED, "reverse mode of a batch handle");
    if (h->P.model == PFDD_CORE_BCC)
        return fail(PFDD_ERR_UNSUPPORTED,
                    "reverse mode supports the sine-series (0, 1, 3, 7, 8), extended (2) and Fourier (4, 9) core models");
    const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, h->tq[3]) : tile_lines_rt<float>(h->nz, h->tq[3]);
    if (((long long)h->nx * h->ny) % nl != 0)
        return fail(PFDD_ERR_UNSUPPORTED, "reverse mode needs nx*ny divisible by %d", nl);
    pfdd_gl::Tape &T = h->tape;
    T.valid = false;
    const size_t fb = (size_t)h->SM * h->N * h->esz;
    const size_t nck = (size_t)((nsteps + every - 1) / every);
    while (T.ckpt.size() > nck) {
        HIPCHK(hipFree(T.ckpt.back()));
        T.ckpt.pop_back();
    }
    while (T.ckpt.size() < nck) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, fb));
        T.ckpt.push_back(p);
    }
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
    HIPCHK(hipGetLastError());
    const int cur0 = h->cur;
    for (int s = 0; s < nsteps; ++s) {
        if (s % every == 0)
            HIPCHK(hipMemcpyAsync(T.ckpt[s / every], h->xi[h->cur], fb, hipMemcpyDeviceToDevice, h->stream));
        rc = h->prec == PFDD_FP64 ? step_once<double>(h, 0.0, nullptr) : step_once<float>(h, 0.0, nullptr);
        if (rc) return rc;
    }
    pfdd_gl_reduce r{};
    if ((rc = collect_steps(h, cur0, &r))) return rc;
    if (out) *out = r;
    T.nsteps = nsteps;
    T.every = every;
    T.P = h->P;
    T.valid = true;
    return PFDD_OK;
}

int pfdd_gl_tape_backward(pfdd_gl *h, const double *dxi_T, double *dxi_0, double *dtau, double *dcoef)
{
    return tape_backward_any(h, dxi_T, false, dxi_0, false, dtau, dcoef);
}

int pfdd_gl_tape_backward_device(pfdd_gl *h, const void *dxi_T, void *dxi_0, double *dtau, double *dcoef)
{
    return tape_backward_any(h, dxi_T, true, dxi_0, true, dtau, dcoef);
}

int pfdd_gl_xi_sums(pfdd_gl *h, double *sum_re)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!sum_re) return fail(PFDD_ERR_ARG, "null output");
    const int nb = 256;
    double *part = nullptr;
    HIPCHK(hipMalloc(&part, sizeof(double) * nb * h->SM));
    if (h->prec == PFDD_FP64)
        hipLaunchKernelGGL((k_sys_sums<double>), dim3(nb, h->SM), dim3(NTHR), 0, h->stream,
                           (const double2 *)h->xi[h->cur], h->N, part);
    else
        hipLaunchKernelGGL((k_sys_sums<float>), dim3(nb, h->SM), dim3(NTHR), 0, h->stream,
                           (const float2 *)h->xi[h->cur], h->N, part);
    HIPCHK(hipGetLastError());
    std::vector<double> hp((size_t)nb * h->SM);
    HIPCHK(hipMemcpyAsync(hp.data(), part, sizeof(double) * nb * h->SM, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipFree(part));
    for (int a = 0; a < h->SM; ++a) {
        double s = 0.0, c = 0.0;
        for (int b = 0; b < nb; ++b) {
            const double y = hp[(size_t)a * nb + b] - c;
            const double t = s + y;
            c = (t - s) - y;
            s = t;
        }
        sum_re[a] = s;
    }
    return PFDD_OK;
}


int pfdd_gl_diag_fields(pfdd_gl *h, const double *sigma_rot, double ave_eps[9], double ave_sigma[9],
                        double *data_eps, double *data_sigma)
{
    int rc;
    if ((rc = check_handle(h))) return rc;
    if (!sigma_rot || !ave_eps || !ave_sigma) return fail(PFDD_ERR_ARG, "null argument");
    if (h->M > 1) return fail(PFDD_ERR_UNSUPPORTED, "output-step diagnostics of a batch handle");
    std::vector<double> sums(h->S);
    if ((rc = pfdd_gl_xi_sums(h, sums.data()))) return rc;
    const double nsize = (double)h->N;
    // stressfree_strain (fft_fftw_slab.cpp:3252-3326): ave_epsd, avepst, avepsts
    double ave_epsd[9] = {0}, avepst[9] = {0}, avepsts[9] = {0};
    for (int a = 0; a < h->S; ++a)
        for (int q = 0; q < 9; ++q) ave_epsd[q] += h->eps[a * 9 + q] * sums[a];
    for (int q = 0; q < 9; ++q) ave_epsd[q] /= nsize;
    auto Sijkl = [&](int i, int j, int k, int l) {
        return h->S44 / 4 * ((i == k) * (j == l) + (i == l) * (j == k)) + h->S12 * (i == j) * (k == l);
    };
    auto Cijkl = [&](int i, int j, int k, int l) {
        return h->C44 * ((i == k) * (j == l) + (i == l) * (j == k)) + h->C12 * (i == j) * (k == l);
    };
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    avepst[3 * i + j] += Sijkl(i, j, k, l) * sigma_rot[3 * k + l] * h->mu;
                    avepsts[3 * i + j] += Sijkl(i, j, k, l) * sigma_rot[3 * k + l] * h->mu;
                }
            avepst[3 * i + j] += ave_epsd[3 * i + j];
        }
    // strain() + total_average_strain(): the k-space strain field has zero mean
    // (F(0) = 0), so the average total strain is the symmetric part of avepst.
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ave_eps[3 * i + j] = (avepst[3 * i + j] + avepst[3 * j + i]) / 2.0;
    // stress() (:3474-3553): mean of C:(strain - eps_d - avepsts) + sigma_rot mu
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m)
                for (int l = 0; l < 3; ++l)
                    s += Cijkl(i, j, m, l) * (avepst[3 * m + l] - ave_epsd[3 * m + l] - avepsts[3 * m + l]);
            ave_sigma[3 * i + j] = s + sigma_rot[3 * i + j] * h->mu;
        }
    if (data_eps || data_sigma) {
        if ((rc = set_device(h))) return rc;
        rc = h->prec == PFDD_FP64 ? diag_fields<double>(h, sigma_rot, avepst, avepsts, data_eps, data_sigma)
                                  : diag_fields<float>(h, sigma_rot, avepst, avepsts, data_eps, data_sigma);
        if (rc) return rc;
    }
    return PFDD_OK;
}

int pfdd_gl_diag(pfdd_gl *h, const double *sigma_rot, double ave_eps[9], double ave_sigma[9])
{
    return pfdd_gl_diag_fields(h, sigma_rot, ave_eps, ave_sigma, nullptr, nullptr);
}

int pfdd_gl_profile(pfdd_gl *h, int nsteps, double *ms, int *nkernels)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!ms || nsteps <= 0) return fail(PFDD_ERR_ARG, "need ms[7] and nsteps > 0");
    // all iterations back to back (as in a production run), one wait at the end
    std::vector<hipEvent_t> ev((size_t)12 * nsteps);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
    for (int s = 0; s < nsteps && !rc; ++s)
        rc = h->prec == PFDD_FP64 ? step_once<double>(h, 0.0, &ev[12 * s]) : step_once<float>(h, 0.0, &ev[12 * s]);
    if (!rc && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(PFDD_ERR_HIP, "profile wait failed");
    for (int s = 0; s < nsteps && !rc; ++s) {
        const hipEvent_t *e = &ev[12 * s];
        for (int q = 0; q < 6; ++q) {
            float t = 0;
            hipEventElapsedTime(&t, e[2 * q], e[2 * q + 1]);
            acc[q] += t;
        }
        float t = 0;
        hipEventElapsedTime(&t, e[0], e[11]);
        acc[6] += t;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int q = 0; q < 7; ++q) ms[q] = acc[q] / nsteps;
    if (nkernels) *nkernels = 6;
    return PFDD_OK;
}

int pfdd_gl_traffic(pfdd_gl *h, double *bytes_per_kernel, double *bytes_per_step)
{
    int rc;
    if ((rc = check_handle(h))) return rc;
    const double f = (double)h->SM * h->N * h->esz;
    const double bsz =
        h->stored_B ? (double)h->S * (h->S + 1) / 2 * h->N * (h->prec == PFDD_FP64 ? 8 : 4) : 0.0;
    const double k[5] = {2 * f, 2 * f, 2 * f + bsz, 2 * f, 3 * f};
    double tot = 0;
    for (int q = 0; q < 5; ++q) {
        if (bytes_per_kernel) bytes_per_kernel[q] = k[q];
        tot += k[q];
    }
    if (bytes_per_step) *bytes_per_step = tot;
    return PFDD_OK;
}

} // extern "C"

// ================================================================== graphs
// A hipGraph holding `iters` (even) GL iterations of several handles, replayed
// as one launch.  Kernel arguments are frozen at capture: the graph is
// captured again when a handle's kernel arguments changed (set_tau of a
// single-run handle, set_core_coef), when its current ping-pong buffer is not
// the one captured, or when tol changes.  A batch handle reads tau from
// device memory, so set_tau needs no new capture.
struct pfdd_gl_graph {
    std::vector<pfdd_gl *> hs;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned> ver;
    std::vector<int> cur;
    int iters = 2;
    double tol = 0.0;
};

namespace {

void graph_release(pfdd_gl_graph *g)
{
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
}

int graph_capture(pfdd_gl_graph *g, double tol)
{
    graph_release(g);
    std::vector<hipStream_t> saved(g->hs.size());
    for (size_t i = 0; i < g->hs.size(); ++i) {
        saved[i] = g->hs[i]->stream;
        g->hs[i]->stream = g->stream;
    }
    int rc = PFDD_OK;
    hipError_t e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = fail(PFDD_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    } else {
        for (int it = 0; it < g->iters && !rc; ++it)
            for (pfdd_gl *h : g->hs) {
                rc = h->prec == PFDD_FP64 ? step_once<double>(h, tol, nullptr) : step_once<float>(h, tol, nullptr);
                if (rc) break;
            }
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(g->stream, &graph);
        if (!rc && e != hipSuccess) rc = fail(PFDD_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        g->graph = graph;
    }
    for (size_t i = 0; i < g->hs.size(); ++i) g->hs[i]->stream = saved[i];
    if (rc) {
        graph_release(g);
        return rc;
    }
    HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->tol = tol;
    for (size_t i = 0; i < g->hs.size(); ++i) {
        g->ver[i] = g->hs[i]->version;
        g->cur[i] = g->hs[i]->cur;
    }
    return PFDD_OK;
}

} // namespace

extern "C" {

int pfdd_gl_graph_create(pfdd_gl *const *handles, int n, int iters, pfdd_gl_graph **out)
{
    g_err.clear();
    if (!handles || n < 1 || !out) return fail(PFDD_ERR_ARG, "need n >= 1 handles and an output");
    if (iters < 2 || iters % 2) return fail(PFDD_ERR_ARG, "iters must be even and >= 2");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!handles[i]) return fail(PFDD_ERR_ARG, "null handle %d", i);
        if (handles[i]->device != handles[0]->device) return fail(PFDD_ERR_ARG, "handles on different devices");
    }
    pfdd_gl_graph *g = new pfdd_gl_graph();
    g->hs.assign(handles, handles + n);
    g->iters = iters;
    g->ver.assign(n, 0);
    g->cur.assign(n, -1);
    hipError_t e = hipSetDevice(handles[0]->device);
    if (e == hipSuccess) e = hipStreamCreate(&g->stream);
    if (e != hipSuccess) {
        delete g;
        return fail(PFDD_ERR_HIP, "graph stream: %s", hipGetErrorString(e));
    }
    *out = g;
    return PFDD_OK;
}

int pfdd_gl_graph_run(pfdd_gl_graph *g, int nsteps, double tol, pfdd_gl_reduce *out)
{
    if (!g) return fail(PFDD_ERR_ARG, "null graph");
    if (nsteps < 0) return fail(PFDD_ERR_ARG, "nsteps < 0");
    HIPCHK(hipSetDevice(g->hs[0]->device));
    const size_t n = g->hs.size();
    bool stale = !g->exec || tol != g->tol;
    for (size_t i = 0; i < n; ++i) stale |= g->hs[i]->version != g->ver[i] || g->hs[i]->cur != g->cur[i];
    // order after each handle's earlier work
    for (pfdd_gl *h : g->hs) HIPCHK(hipStreamSynchronize(h->stream));
    int rc;
    if (stale && (rc = graph_capture(g, tol))) return rc;
    std::vector<int> cur0(n);
    for (size_t i = 0; i < n; ++i) {
        pfdd_gl *h = g->hs[i];
        cur0[i] = h->cur;
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, g->stream, h->done, h->steps);
        HIPCHK(hipGetLastError());
    }
    for (int r = 0; r < nsteps / g->iters; ++r) HIPCHK(hipGraphLaunch(g->exec, g->stream));
    // the remainder eagerly, on the graph stream
    const int rem = nsteps % g->iters;
    for (int it = 0; it < rem; ++it)
        for (pfdd_gl *h : g->hs) {
            hipStream_t s0 = h->stream;
            h->stream = g->stream;
            rc = h->prec == PFDD_FP64 ? step_once<double>(h, tol, nullptr) : step_once<float>(h, tol, nullptr);
            h->stream = s0;
            if (rc) return rc;
        }
    HIPCHK(hipStreamSynchronize(g->stream));
    int first_err = PFDD_OK;
    for (size_t i = 0; i < n; ++i) {
        pfdd_gl_reduce r{};
        rc = collect_steps(g->hs[i], cur0[i], &r);
        if (rc && !first_err) first_err = rc;
        if (out) out[i] = r;
    }
    return first_err;
}

int pfdd_gl_graph_destroy(pfdd_gl_graph *g)
{
    if (!g) return PFDD_OK;
    hipSetDevice(g->hs[0]->device);
    if (g->stream) hipStreamSynchronize(g->stream);
    graph_release(g);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return PFDD_OK;
}

} // extern "C"

// ============================================================ conductivity
// The conductivity sub-problem of solve_style GLCond (solve_GL+cond.cpp:
// 70-150): one scalar complex field c (fft->condxi) on the same grid, stepped
// by forwardcond (fft_fftw_slab.cpp:1999-2006), divcurrent (:1704-1741),
// backwardcond (:2057-2191) and update_conduct_order_param (:1794-1835).
// The layer conductivities (backwardcond's per-row material algebra) come in
// from the host as two arrays over the y index j: sig[j] (condsigma of row
// j, scattering included) and dsig[j] (the difference of the bulk
// conductivity across row j that multiplies the applied-field term).
namespace {

struct CondParams {
    double four[9];
    double El;       // material->El[1]
    double eps0;     // material->epsilonzero
    double rate;     // 1e-22 * app->timestep
    double invN;     // 1 / (nx ny nz)
    int nx, ny, nz;
};

// divcurrent on the transformed field in the [k][j][i] layout: a = i c f_y /
// f^2 and b = c f_y^2 / f^2 (zero at f = 0), written as [2][k][j][i].  The
// block also sums sig[(ny - j) mod ny] Re a over its k-points: backwardcond
// called a second time on its own output (the trailing call of
// SolveGLCond::iterate) transforms a back once more, and two unnormalised
// inverse transforms give N a(-r).
__global__ void __launch_bounds__(NTHR)
    k_cond_div(const double2 *__restrict__ ch, double2 *__restrict__ ab, const double *__restrict__ sig, CondParams P,
               long long N, double *__restrict__ part)
{
    __shared__ double red[NTHR / 64];
    const int tid = threadIdx.x;
    const long long idx = (long long)blockIdx.x * NTHR + tid;
    double v[1] = {0.0};
    if (idx < N) {
        const int i = (int)(idx % P.nx);
        const int j = (int)((idx / P.nx) % P.ny);
        const int k = (int)(idx / ((long long)P.nx * P.ny));
        const double kx = frac_freq(i, P.nx), ky = frac_freq(j, P.ny), kz = frac_freq(k, P.nz);
        double f[3];
        for (int c = 0; c < 3; ++c) f[c] = kx * P.four[c] + ky * P.four[3 + c] + kz * P.four[6 + c];
        const double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        const double2 c = ch[idx];
        double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
        if (f2 > 0.0) {
            a.x = -(c.y * f[1] / f2);
            a.y = c.x * f[1] / f2;
            b.x = c.x * f[1] * f[1] / f2;
            b.y = c.y * f[1] * f[1] / f2;
        }
        ab[idx] = a;
        ab[N + idx] = b;
        v[0] = sig[(P.ny - j) % P.ny] * a.x;
    }
    block_sum<1, NTHR>(v, red, tid);
    if (tid == 0) part[blockIdx.x] = v[0];
}

// backwardcond's per-cell assembly on the (unnormalised) inverse transforms
// a, b in the [2][i][j][k] layout, then update_conduct_order_param:
//   j     = sig (El - Re a / eps0) - i sig Im a / eps0
//   rhs   = (sig / eps0) b + (a / eps0 - El) dsig      (El on the real part)
//   c    -= rate * rhs / N
// Block sums: Re j, Im j, |dc|^2, Re c_new.
__global__ void __launch_bounds__(NTHR)
    k_cond_update(const double2 *__restrict__ ab, double2 *__restrict__ c, const double *__restrict__ sig,
                  const double *__restrict__ dsig, CondParams P, long long N, double *__restrict__ part,
                  long long pstride)
{
    __shared__ double red[4 * (NTHR / 64)];
    const int tid = threadIdx.x;
    const long long idx = (long long)blockIdx.x * NTHR + tid;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (idx < N) {
        const int j = (int)((idx / P.nz) % P.ny);
        const double2 a = ab[idx], b = ab[N + idx];
        const double s = sig[j], ds = dsig[j];
        const double jr = s * (P.El - a.x / P.eps0);
        const double ji = -s * a.y / P.eps0;
        double rr = (s / P.eps0) * b.x, ri = (s / P.eps0) * b.y;
        rr += (a.x / P.eps0 - P.El) * ds;
        ri += (a.y / P.eps0) * ds;
        const double2 c0 = c[idx];
        double2 c1;
        c1.x = c0.x - (P.rate * (rr * P.invN));
        c1.y = c0.y - (P.rate * (ri * P.invN));
        c[idx] = c1;
        const double dr = c1.x - c0.x, di = c1.y - c0.y;
        v[0] = jr;
        v[1] = ji;
        v[2] = dr * dr + di * di;
        v[3] = c1.x;
    }
    block_sum<4, NTHR>(v, red, tid);
    if (tid == 0)
        for (int q = 0; q < 4; ++q) part[q * pstride + blockIdx.x] = v[q];
}

} // namespace

struct pfdd_gl_cond {
    int device = 0, nx = 0, ny = 0, nz = 0;
    long long N = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    CondParams P{};
    void *tw[3] = {nullptr, nullptr, nullptr};
    double2 *c = nullptr, *w0 = nullptr, *w1 = nullptr;
    double *sig = nullptr, *dsig = nullptr, *part = nullptr;
    int *done = nullptr;
    std::vector<double> sig_h;
    long long nb = 0;
};

namespace {

void cond_release(pfdd_gl_cond *h)
{
    for (void *p : {(void *)h->c, (void *)h->w0, (void *)h->w1, (void *)h->sig, (void *)h->dsig, (void *)h->part,
                    (void *)h->done, h->tw[0], h->tw[1], h->tw[2]})
        if (p) hipFree(p);
    if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
}

// Same line maps as the GL passes, for S fields of the cond grid.
Maps cond_maps(const pfdd_gl_cond *h, long long S) { return make_maps_dims(h->nx, h->ny, h->nz, S); }

int cond_step(pfdd_gl_cond *h, pfdd_gl_cond_reduce *out)
{
    hipStream_t st = h->stream;
    const Maps m1 = cond_maps(h, 1), m2 = cond_maps(h, 2);
    const long long ncol = (long long)h->ny * h->nz;
    const LineMap xs = {h->nx, 0, 0, 1, 1, 1};
    // forwardcond: FFT of c (z, y, x passes) -> w0 in [k][j][i]
    HIPCHK((launch_pass<double, false>(h->nz, h->c, h->w0, m1.p1i, m1.p1o, m1.n1, h->tw[2], h->done, st)));
    HIPCHK((launch_pass<double, false>(h->ny, h->w0, h->w1, m1.p2i, m1.p2o, m1.n2, h->tw[1], h->done, st)));
    HIPCHK((launch_pass<double, false>(h->nx, h->w1, h->w0, xs, xs, ncol, h->tw[0], h->done, st)));
    // divcurrent -> w1 = [a, b] in k space
    const unsigned nb = (unsigned)h->nb;
    hipLaunchKernelGGL(k_cond_div, dim3(nb), dim3(NTHR), 0, st, (const double2 *)h->w0, h->w1, h->sig, h->P, h->N,
                       h->part + 4 * h->nb);
    HIPCHK(hipGetLastError());
    // backwardcond's two inverse transforms -> w0 = [a, b] in [2][i][j][k]
    HIPCHK((launch_pass<double, true>(h->nx, h->w1, h->w0, xs, xs, 2 * ncol, h->tw[0], h->done, st)));
    HIPCHK((launch_pass<double, true>(h->ny, h->w0, h->w1, m2.p2o, m2.p2i, m2.n2, h->tw[1], h->done, st)));
    HIPCHK((launch_pass<double, true>(h->nz, h->w1, h->w0, m2.p5i, m2.p5o, m2.n5, h->tw[2], h->done, st)));
    hipLaunchKernelGGL(k_cond_update, dim3(nb), dim3(NTHR), 0, st, (const double2 *)h->w0, h->c, h->sig, h->dsig,
                       h->P, h->N, h->part, h->nb);
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)5 * h->nb);
    HIPCHK(hipMemcpyAsync(part.data(), h->part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // block partials summed in block order: deterministic
    double s[5] = {0, 0, 0, 0, 0};
    for (int q = 0; q < 5; ++q)
        for (long long b = 0; b < h->nb; ++b) s[q] += part[(size_t)q * h->nb + b];
    if (!std::isfinite(s[0]) || !std::isfinite(s[2]) || !std::isfinite(s[3]))
        return fail(PFDD_ERR_NONFINITE, "non-finite conductivity reduction");
    if (out) {
        double mean_sig = 0.0;
        for (double v : h->sig_h) mean_sig += v;
        mean_sig /= (double)h->ny;
        out->javerage_re = s[0] * h->P.invN;
        out->javerage_im = s[1] * h->P.invN;
        out->conductivity = out->javerage_re / h->P.El;
        out->condxinorm = std::sqrt(s[2]);
        out->condxiave = s[3];
        out->conductivity_repeat = (h->P.El * mean_sig - s[4] / h->P.eps0) / h->P.El;
    }
    return PFDD_OK;
}

} // namespace

extern "C" {

int pfdd_gl_cond_create(const pfdd_gl_cond_desc *d, pfdd_gl_cond **out)
{
    if (!d || !out) return fail(PFDD_ERR_ARG, "null argument");
    *out = nullptr;
    if (!size_supported(d->nx) || !size_supported(d->ny) || !size_supported(d->nz))
        return fail(PFDD_ERR_UNSUPPORTED, "grid %dx%dx%d: each size must be one of 8 16 32 40 48 64 80 96 128 160 192 256 512",
                    d->nx, d->ny, d->nz);
    if (!d->sigma || !d->dsigma) return fail(PFDD_ERR_ARG, "null layer conductivity arrays");
    if (!(d->El != 0.0) || !(d->epsilonzero > 0.0))
        return fail(PFDD_ERR_ARG, "need a nonzero applied field El and epsilonzero > 0");
    pfdd_gl_cond *h = new pfdd_gl_cond;
    auto bail = [&](int rc) {
        cond_release(h);
        delete h;
        return rc;
    };
    h->device = d->device;
    h->nx = d->nx;
    h->ny = d->ny;
    h->nz = d->nz;
    h->N = (long long)d->nx * d->ny * d->nz;
    h->nb = (h->N + NTHR - 1) / NTHR;
    for (int q = 0; q < 9; ++q) h->P.four[q] = d->four[q];
    h->P.El = d->El;
    h->P.eps0 = d->epsilonzero;
    h->P.rate = d->rate;
    h->P.invN = 1.0 / (double)h->N;
    h->P.nx = d->nx;
    h->P.ny = d->ny;
    h->P.nz = d->nz;
    h->sig_h.assign(d->sigma, d->sigma + d->ny);
    hipError_t e = hipSetDevice(d->device);
    if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e)));
    if (d->stream) {
        h->stream = (hipStream_t)d->stream;
    } else {
        e = hipStreamCreate(&h->stream);
        if (e != hipSuccess) return bail(fail(PFDD_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
        h->own_stream = true;
    }
    const size_t fb = sizeof(double2) * (size_t)h->N;
    if ((e = hipMalloc((void **)&h->c, fb)) != hipSuccess || (e = hipMalloc((void **)&h->w0, 2 * fb)) != hipSuccess ||
        (e = hipMalloc((void **)&h->w1, 2 * fb)) != hipSuccess ||
        (e = hipMalloc((void **)&h->sig, sizeof(double) * d->ny)) != hipSuccess ||
        (e = hipMalloc((void **)&h->dsig, sizeof(double) * d->ny)) != hipSuccess ||
        (e = hipMalloc((void **)&h->part, sizeof(double) * 5 * h->nb)) != hipSuccess ||
        (e = hipMalloc((void **)&h->done, sizeof(int))) != hipSuccess)
        return bail(fail(PFDD_ERR_HIP, "hipMalloc: %s", hipGetErrorString(e)));
    const int dims[3] = {h->nx, h->ny, h->nz};
    for (int q = 0; q < 3; ++q) {
        std::vector<double> tw;
        make_twiddles(dims[q], tw);
        if ((e = hipMalloc(&h->tw[q], sizeof(double2) * dims[q])) != hipSuccess ||
            (e = hipMemcpy(h->tw[q], tw.data(), sizeof(double2) * dims[q], hipMemcpyHostToDevice)) != hipSuccess)
            return bail(fail(PFDD_ERR_HIP, "twiddles: %s", hipGetErrorString(e)));
    }
    const int zero = 0;
    if ((e = hipMemcpy(h->sig, d->sigma, sizeof(double) * d->ny, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->dsig, d->dsigma, sizeof(double) * d->ny, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->done, &zero, sizeof(int), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(h->c, 0, fb)) != hipSuccess)
        return bail(fail(PFDD_ERR_HIP, "upload: %s", hipGetErrorString(e)));
    *out = h;
    return PFDD_OK;
}

int pfdd_gl_cond_destroy(pfdd_gl_cond *h)
{
    if (!h) return PFDD_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    cond_release(h);
    delete h;
    return PFDD_OK;
}

int pfdd_gl_cond_set(pfdd_gl_cond *h, const double *c_host)
{
    if (!h || !c_host) return fail(PFDD_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(h->c, c_host, sizeof(double2) * h->N, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return PFDD_OK;
}

int pfdd_gl_cond_get(pfdd_gl_cond *h, double *c_host)
{
    if (!h || !c_host) return fail(PFDD_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(h->device));
    HIPCHK(hipMemcpyAsync(c_host, h->c, sizeof(double2) * h->N, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return PFDD_OK;
}

int pfdd_gl_cond_step(pfdd_gl_cond *h, int nsteps, pfdd_gl_cond_reduce *out)
{
    if (!h) return fail(PFDD_ERR_ARG, "null handle");
    if (nsteps < 1) return fail(PFDD_ERR_ARG, "nsteps must be >= 1");
    HIPCHK(hipSetDevice(h->device));
    for (int s = 0; s < nsteps; ++s) {
        const int rc = cond_step(h, out);
        if (rc) return rc;
    }
    return PFDD_OK;
}

} // extern "C"
