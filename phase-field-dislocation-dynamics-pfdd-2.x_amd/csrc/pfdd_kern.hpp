// pfdd_kern.hpp -- declarations shared by the translation units of the GL
// step library: constants, kernel parameter blocks, core-energy device
// functions, the grid-size dispatch table and the kernel launchers (defined
// in pfdd_pass.hip, pfdd_update.hip, pfdd_mid.hip and pfdd_plane.hip).
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "pfdd_fft.hpp"
#include "../../include/pfdd_gl.h"

namespace pfdd {

constexpr int MAXS = PFDD_GL_MAX_SYSTEMS;
constexpr int MAXM = PFDD_GL_MAX_COEF;
constexpr int NTHR = 256;
constexpr int NSLOT = 5; // xinorm^2, xiave, e_core, e_stress, e_el
constexpr int NRES = 8;  // xinorm, xiave, e_el, e_core, e_stress, e_total, e_core_ref, nonfinite

// Kernel-timing events for the next launch of the GL iteration's kernels
// (set by the profiler around one launch at a time): the dispatch itself
// records them (hipExtLaunchKernelGGL), so they bracket exactly the kernel.
extern thread_local hipEvent_t t_ev_start, t_ev_stop;

#define PFDD_LAUNCH(kern, grid, block, st, ...)                                                        \
    do {                                                                                              \
        if (t_ev_start)                                                                               \
            hipExtLaunchKernelGGL(kern, grid, block, 0, st, t_ev_start, t_ev_stop, 0, __VA_ARGS__);    \
        else                                                                                          \
            hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__);                                \
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

// sin(pi x) and cos(pi x) for the core-energy derivatives.  x = n + r with
// n = rint(x) and r = x - n (exact, |r| <= 1/2); sin(pi r) = r P(r^2) and
// cos(pi r) = Q(r^2), then both times (-1)^n.  P and Q interpolate
// sin(pi r)/r and cos(pi r) at the Chebyshev points of r^2 in [0, 1/4]
// (degree 8 in FP64, in long double, then rounded; degree 4 in FP32, a
// least-squares fit on 4000 Chebyshev points with relative weights for sin).
// Evaluated by Horner with FMA: at most 1.3 ulp (FP64) and 1.4 ulp (FP32)
// from sin/cos(pi x) on |x| <= 1/2 (checked against long double on 4e5
// points), against ~45 (FP32) and more (FP64) VALU instructions of the
// library's sincospi, which the VALU-bound update kernels feel.  NaN and
// infinite x give NaN.
__device__ __forceinline__ void sincospi_poly(double x, double &s, double &c)
{
    const double n = __builtin_rint(x);
    const double r = x - n, t = r * r;
    double p = 0x1.9d2aaaaaaaaabp-21;
    p = __builtin_fma(p, t, -0x1.6f79d8e38e38ep-16);
    p = __builtin_fma(p, t, 0x1.e8f35feaaaaabp-12);
    p = __builtin_fma(p, t, -0x1.e3074ddaa0000p-8);
    p = __builtin_fma(p, t, 0x1.50783484ff78ep-4);
    p = __builtin_fma(p, t, -0x1.32d2cce626dd4p-1);
    p = __builtin_fma(p, t, 0x1.466bc6775aa52p+1);
    p = __builtin_fma(p, t, -0x1.4abbce625be52p+2);
    p = __builtin_fma(p, t, 0x1.921fb54442d18p+1);
    double q = 0x1.167771c71c71cp-18;
    q = __builtin_fma(q, t, -0x1.b69562aaaaaabp-14);
    q = __builtin_fma(q, t, 0x1.f9d253ac38e39p-10);
    q = __builtin_fma(q, t, -0x1.a6d1efc621c72p-6);
    q = __builtin_fma(q, t, 0x1.e1f506812dcb9p-3);
    q = __builtin_fma(q, t, -0x1.55d3c7e3bf9d1p+0);
    q = __builtin_fma(q, t, 0x1.03c1f081b598cp+2);
    q = __builtin_fma(q, t, -0x1.3bd3cc9be45dbp+2);
    q = __builtin_fma(q, t, 1.0);
    // (-1)^n from the parity of the integer-valued n (exact for |n| < 2^53)
    const double h = 0.5 * n;
    const bool odd = h != __builtin_rint(h);
    s = odd ? -(r * p) : r * p;
    c = odd ? -q : q;
}
__device__ __forceinline__ void sincospi_poly(float x, float &s, float &c)
{
    const float n = __builtin_rintf(x);
    const float r = x - n, t = r * r;
    float p = 0x1.3e13b0p-4f;
    p = __builtin_fmaf(p, t, -0x1.3253a8p-1f);
    p = __builtin_fmaf(p, t, 0x1.4668fep+1f);
    p = __builtin_fmaf(p, t, -0x1.4abbc4p+2f);
    p = __builtin_fmaf(p, t, 0x1.921fb6p+1f);
    float q = 0x1.c29b9cp-3f;
    q = __builtin_fmaf(q, t, -0x1.550192p+0f);
    q = __builtin_fmaf(q, t, 0x1.03bd86p+2f);
    q = __builtin_fmaf(q, t, -0x1.3bd3aep+2f);
    q = __builtin_fmaf(q, t, 1.0f);
    const float h = 0.5f * n;
    const bool odd = h != __builtin_rintf(h);
    s = odd ? -(r * p) : r * p;
    c = odd ? -q : q;
}

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
        sincospi_poly((R)2 * arg[q], sn[q], cs[q]);
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
        sincospi_poly((R)2 * arg, sn, cs);
        const bool is_sin = G[t] >= 5;
        const double Tv = is_sin ? (double)sn : (double)cs;
        const double Td = is_sin ? (double)cs : -(double)sn;
        const double uk = (double)U[t][k];
        h -= tp * tp * (double)c[G[t]] * uk * (double)um * Tv;
        gc[G[t]] += (double)mu[k] * tp * uk * Td;
    }
    hv = (R)h;
}

// One cell of update_order_parameter (fft_fftw_slab.cpp:1746-1790): the
// core-energy derivative of core_energy_* at Re xi of the old field x (the
// reference evaluates it on data_fftw = xi before the update), the applied
// resolved shear tau and the elastic force g (unnormalised inverse transform
// of ghat, hence invN).  Masked cells keep their value; their xinorm term
// uses the previous unmasked cell's old value (quirk Q3, prev[] built on the
// host, -1 = the initial 0).  coef(m) is the cell's m-th core coefficient,
// sub the position of its system in the plane's triple (coupled model).
// Returns the new value; dn is the cell's |xi_new - xi_old|^2 term of xinorm
// and dens its core-energy density.
template <typename R, bool GEN, typename V, typename Coef>
__device__ __forceinline__ V gl_update_cell(const V g, const V x, long long idx, R tau, int sub, Coef coef,
                                            const StepParams &P, int nm, R cddt, R invN,
                                            const uint8_t *__restrict__ mask, const long long *__restrict__ prev,
                                            const V *__restrict__ xo, const R *__restrict__ amp, double &dn, R &dens)
{
    R dE = (R)0;
    dens = (R)0;
    if (GEN && P.model == PFDD_CORE_EXTENDED) {
        // the other two systems of the plane at the same cell
        const int j = sub;
        const long long b0 = idx - (long long)j * P.ncell;
        R y[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) y[q] = q == j ? x.x : xo[b0 + q * P.ncell].x;
        R cc[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) cc[q] = coef(q);
        R f;
        core_extended<R>(y[j], y[(j + 1) % 3], y[(j + 2) % 3], cc, f, dE);
        dens = f / (R)3; // each of the three systems carries a third of the plane's density
    } else {
        R sp, cp;
        sincospi_poly(x.x, sp, cp);
        if (GEN && (P.model == PFDD_CORE_PYRII || P.model == PFDD_CORE_FOURIER)) {
            // a0 + sum_m a_m cos(2 pi m x) + b_m sin(2 pi m x), m <= 4
            const R s1 = (R)2 * sp * cp, c1 = cp * cp - sp * sp;
            R sm = s1, cm = c1;
            dens = (R)P.a0;
#pragma unroll
            for (int m = 1; m <= 4; ++m) {
                const R am = coef(m - 1), bm = coef(m + 3);
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
                const R c = coef(m - 1);
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
    V y;
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
    return y;
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

// ------------------------------------------------------------------ dispatch
#define PFDD_SIZES(X) X(8) X(16) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(512)

inline bool size_supported(int n)
{
#define PFDD_CASE(n_) if (n == n_) return true;
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return false;
}

// Whether the line passes have a distinct radix-16 plan for length n.
inline bool radix16_plan_differs_rt(int n)
{
#define PFDD_CASE(n_) if (n == n_) return radix16_plan_differs<n_>();
    PFDD_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return false;
}

inline bool systems_supported(int s) { return s == 1 || s == 2 || s == 3 || s == 4 || s == 6 || s == 12; }

// Lines per tile of the update kernels for tile shape tq.
template <typename R>
int tile_lines_rt(int n, int tq)
{
    tq %= 3;
#define PFDD_CASE(n_)                                              \
    if (n == n_) {                                                 \
        if (tq == 0) return tile_lines_u<R, n_, 0>();              \
        if (tq == 2) return tile_lines_u<R, n_, 2>();              \
        return tile_lines_u<R, n_, 1>();                           \
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

// ---------------------------------------------------------------- launchers
// Each returns hipGetLastError() of its launch, or hipErrorInvalidValue for a
// grid size (or system count) outside the dispatch table.

// Line pass (pfdd_pass.hip): FFT along one axis of nlines lines, input line
// map mi, output line map mo; zmask (optional) zeroes input elements.  tq =
// tile shape (0, 1, 2) + 3 * radix plan (0 radix 8 first, 1 radix 16 where it
// saves a stage).
template <typename R, bool INV>
hipError_t launch_pass(int n, const void *in, void *out, LineMap mi, LineMap mo, long long nlines,
                       const void *tw, const int *done, hipStream_t st, const uint8_t *zmask = nullptr,
                       int tq = 1);

// Last inverse pass + GL update, and its reverse mode (pfdd_update.hip).
template <typename R>
hipError_t launch_update(int n, int tq, const void *T1, const void *xo, void *xn, const uint8_t *mask,
                         const long long *prev, const void *amp, LineMap mi, LineMap mo, long long nlines,
                         long long lps, const StepParams &P, const void *tw, double *partial,
                         long long pstride, const int *done, hipStream_t st);
template <typename R>
hipError_t launch_adj_update(int n, int tq, const void *T1, const void *lam, const void *xs, void *lam_out,
                             const uint8_t *mask, const void *amp, LineMap mi, LineMap mo, long long nlines,
                             long long lps, const StepParams &P, const void *tw, double *gpart,
                             long long gstride, const int *done, hipStream_t st);

// x pass (pfdd_mid.hip): blocks per member, support check, stored-B and
// on-the-fly variants.
template <typename R> int mid_blocks(int n, int S, long long ncol);
template <typename R> bool mid_supported(int n, int S);
template <typename R>
hipError_t launch_mid(int n, int S, int M, void *T2, const void *B, long long ncol, long long ntot, const void *tw,
                      double *part_el, const int *done, hipStream_t st);
template <typename R>
hipError_t launch_mid_otf(int n, int S, int M, void *T2, const GParams &G, long long ncol, long long ntot,
                          const void *tw, double *part_el, const int *done, hipStream_t st, int pipe, int ncu,
                          int *nblocks);
template <typename R>
hipError_t launch_mid_pl(int n, int S, int M, void *T, const GParams &G, long long ncol, long long ntot,
                         const void *tw, double *part_el, const int *done, hipStream_t st, int *nblocks);
template <typename R> int mid_pl_blocks(int n, int S, int M, long long ncol);

// Plane iteration (pfdd_plane.hip): the 2-D forward pass over y-z planes and
// the 2-D inverse pass fused with the update, for square y-z planes that fit
// in LDS; persistent grids of plane_blocks blocks.
template <typename R> bool plane_supported(int ny, int nz);
// which: 0 the forward kernel, 1 the update, 2 the update of the coupled /
// Fourier core models.
template <typename R> int plane_blocks(int ny, int nz, long long nplanes, int ncu, int which);
template <typename R>
hipError_t launch_plane_fwd(int ny, int nz, const void *in, void *out, long long nplanes, int nblocks,
                            const void *twy, const void *twz, const int *done, hipStream_t st);
template <typename R>
hipError_t launch_plane_upd(int ny, int nz, const void *T, const void *xo, void *xn, const uint8_t *mask,
                            const long long *prev, const void *amp, long long nplanes, int nx, int nblocks,
                            const StepParams &P, const void *twy, const void *twz, double *partial,
                            long long pstride, const int *done, hipStream_t st);

} // namespace pfdd
