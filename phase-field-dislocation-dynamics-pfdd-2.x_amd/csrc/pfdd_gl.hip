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

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pfdd_kern.hpp"

using namespace pfdd;

namespace pfdd {
thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;
} // namespace pfdd

namespace {

thread_local std::string g_err;

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
// block) into stage2; the last block to finish (atomic ticket) sums stage2
// with all its threads (one value per thread and slot, the compensated block
// tree), writes the iteration's scalars, counts the iteration and raises
// `done` on convergence or a non-finite result.
constexpr int NB2 = 32;
static_assert(NB2 <= NTHR, "the last block reads stage2 one value per thread");

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
    if (!last) return;
    __threadfence();
    // one dependent round trip instead of NB2 * NSLOT serial loads in one
    // thread (the first level's stage2 writes are visible: ticket + fences)
    double w[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) w[s] = tid < NB2 ? __builtin_nontemporal_load(&stage2[s * NB2 + tid]) : 0.0;
    block_sum_comp<NSLOT, NTHR>(w, red, tid);
    if (tid != 0) return;
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

// Streaming rate probes for pfdd_gl_stream_rates: one 16-byte element per
// thread, nontemporal loads and stores (the access pattern the field passes
// use), copy (1 read + 1 write) and add (2 reads + 1 write).
__global__ void __launch_bounds__(256)
    k_stream_copy(const double2 *__restrict__ a, double2 *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) st_stream(&b[i], ld_stream(&a[i]));
}
__global__ void __launch_bounds__(256)
    k_stream_add(const double2 *__restrict__ a, const double2 *__restrict__ c, double2 *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const double2 x = ld_stream(&a[i]), y = ld_stream(&c[i]);
        st_stream(&b[i], make_double2(x.x + y.x, x.y + y.y));
    }
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

// Final stage of the per-system sums: system a = blockIdx.x adds its nb
// block partials in index order with Kahan compensation, the same sequence of
// operations as the host loop of pfdd_gl_xi_sums (additions only, so no
// contraction into FMAs: the two give identical sums).
__global__ void k_sys_sums_final(const double *__restrict__ part, int nb, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    const int a = blockIdx.x;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nb; ++b) {
        const double y = part[(long long)a * nb + b] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    out[a] = s;
}

// Device scratch of pfdd_gl_xi_sums_noalloc: XS_NB block partials per system
// for up to PFDD_GL_NOALLOC_MAX_SYSTEMS systems, then one sum per system.
// Every slot that a call reads is written earlier in the same call, so the
// array needs no clearing.  Host code never uses the symbol itself, only the
// per-device address that noalloc_scratch() resolves.
constexpr int XS_NB = 256;
constexpr long long XS_SCRATCH = (long long)(XS_NB + 1) * PFDD_GL_NOALLOC_MAX_SYSTEMS;
static __device__ __align__(256) double scratch[XS_SCRATCH];

// Address of `scratch` on the current device, resolved once per device with
// hipGetSymbolAddress (each GPU holds its own copy of the array).
int noalloc_scratch(int device, double **ptr)
{
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return PFDD_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(PFDD_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    addr[device] = (double *)p;
    *ptr = (double *)p;
    return PFDD_OK;
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
    int plane = 0;                     // 1: three-kernel plane iteration (pfdd_plane.hip)
    // pfdd_gl_step's own graphs: SG_ITERS iterations captured on the
    // handle's stream, replayed for long runs, one per ping-pong parity
    // (sg[cur] starts from xi[cur]); captured again when the kernel
    // arguments (version) or tol change
    struct SelfGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned ver = 0;
        double tol = 0.0;
    } sg[2];
    // false when the environment sets PFDD_GL_GRAPH=0 at handle creation:
    // every iteration is launched eagerly (for tools that cannot follow
    // graph launches; rocprofv3's kernel trace faults on them here)
    bool sg_on = true;
    bool plane_ok = false;             // the grid and precision allow the plane iteration
    int nbA = 0, nbC = 0, nb3pl = 0;   // grids of the plane iteration's kernels
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
    if (h->plane) {
        // plane iteration: 2-D forward pass xo -> T1, x pass on T1 in its
        // plane layout, 2-D inverse pass + update T1 + xo -> T2.  The slots
        // of the line passes it does not have (1: fwd y, 3: inv y) get empty
        // event pairs, so profiles keep their six-kernel shape.
        const long long nplanes = (long long)h->SM * h->nx;
        auto empty = [&](int k) -> int {
            if (ev) {
                HIPCHK(hipEventRecord(ev[2 * k], st));
                HIPCHK(hipEventRecord(ev[2 * k + 1], st));
            }
            return PFDD_OK;
        };
        int rc;
        arm(0);
        HIPCHK(launch_plane_fwd<R>(h->ny, h->nz, xo, h->T1, nplanes, h->nbA, h->tw[1], h->tw[2], h->done, st));
        t_ev_start = t_ev_stop = nullptr;
        if ((rc = empty(1))) return rc;
        arm(2);
        HIPCHK(launch_mid_pl<R>(h->nx, h->S, h->M, h->T1, h->G, (long long)h->ny * h->nz, h->N, h->tw[0],
                                h->partial + 4 * h->pstride, h->done, st, &h->nb3));
        t_ev_start = t_ev_stop = nullptr;
        if ((rc = empty(3))) return rc;
        arm(4);
        HIPCHK(launch_plane_upd<R>(h->ny, h->nz, h->T1, xo, T2, h->mask, h->prev, h->cof ? h->cof : h->amp, nplanes,
                                   h->nx, h->nbC, P, h->tw[1], h->tw[2], h->partial, h->pstride, h->done, st));
    } else {
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
    }
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
    // tq[4]: 0 / 1 the five-kernel iteration with the plain / pipelined x
    // pass, 2 the three-kernel plane iteration where the grid allows it
    h->plane = (tq[4] == 2 && h->plane_ok) ? 1 : 0;
    h->xpipe = (h->stored_B || tq[4] != 1) ? 0 : 1;
    if (h->plane) {
        h->nb5 = (int)((long long)h->SM * h->nx); // one partial per y-z plane
        h->nb3 = h->nb3pl;
    } else {
        const long long n5 = (long long)h->SM * h->nx * h->ny;
        const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, h->tq[3])
                                            : tile_lines_rt<float>(h->nz, h->tq[3]);
        h->nb5 = (int)((n5 + nl - 1) / nl);
    }
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

int self_graph_ready(pfdd_gl *h, double tol);

struct TuneKey {
    int nx, ny, nz, S, M, prec, flags;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(nx, ny, nz, S, M, prec, flags) < std::tie(o.nx, o.ny, o.nz, o.S, o.M, o.prec, o.flags);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::array<int, 5>> g_tune_cache;

// Times every tile shape (0, 1, 2) of each line pass, with both radix plans
// where the pass length has two (tq 3-5: the shapes with the radix-16 plan),
// and the x-pass / iteration variants, keeps the fastest per kernel and then
// the fastest whole step of that mix and the two best uniform choices.  The
// field, ping-pong parity and step counters are restored.  Choices that would
// break the reverse mode's one-system-per-block rule for the update pass are
// skipped.
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
    const int cur0 = h->cur;
    const bool stepped0 = h->stepped;
    // the current field, and (after a step) the previous one, which the
    // output-step strain of the reference is computed from (Q2)
    const int nsave = stepped0 ? 2 : 1;
    char *save = nullptr;
    HIPCHK(hipMalloc((void **)&save, nsave * fb));
    HIPCHK(hipMemcpyAsync(save, h->xi[cur0], fb, hipMemcpyDeviceToDevice, h->stream));
    if (stepped0) HIPCHK(hipMemcpyAsync(save + fb, h->xi[1 - cur0], fb, hipMemcpyDeviceToDevice, h->stream));
    auto restore = [&]() -> int {
        HIPCHK(hipMemcpyAsync(h->xi[cur0], save, fb, hipMemcpyDeviceToDevice, h->stream));
        if (stepped0) HIPCHK(hipMemcpyAsync(h->xi[1 - cur0], save + fb, fb, hipMemcpyDeviceToDevice, h->stream));
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
        HIPCHK(hipGetLastError());
        h->cur = cur0;
        h->stepped = stepped0;
        return PFDD_OK;
    };
    int arg[5] = {1, 1, 1, 1, 0};
    int rc = PFDD_OK;
    // candidates: tile shape q = 0, 1, 2 with the one-tile-per-block x pass;
    // the same shapes with the radix-16 plan in the line passes (q + 3,
    // where the z or y length has that plan); then (on-the-fly contraction
    // only) the pipelined x pass, then (where the grid allows it) the
    // three-kernel plane iteration.  The clock of an idle GPU ramps up over
    // the first few dozen dispatches, so all candidates are timed after a
    // common warm-up, in two interleaved rounds, keeping each kernel's best
    // time per candidate.  The five-kernel iteration takes the best shape and
    // plan of each kernel; the plane iteration replaces it if its whole step
    // (three kernels) is faster than that sum.
    const bool plan_z = radix16_plan_differs_rt(h->nz), plan_y = radix16_plan_differs_rt(h->ny);
    const int nshape = (plan_z || plan_y) ? 6 : 3;
    const int c_pipe = h->stored_B ? -1 : nshape;
    const int c_plane = (!h->stored_B && h->plane_ok) ? nshape + 1 : -1;
    const int ncand = nshape + (c_pipe >= 0 ? 1 : 0) + (c_plane >= 0 ? 1 : 0);
    double tmin[8][5];
    for (int c = 0; c < 8; ++c)
        for (int p = 0; p < 5; ++p) tmin[c][p] = 1e300;
    {
        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
        double warm[5];
        rc = time_passes<R>(h, 20, warm);
    }
    for (int round = 0; round < 2 && !rc; ++round)
        for (int c = 0; c < ncand && !rc; ++c) {
            const int ln = c < nshape ? c : 1, up = c < nshape ? c % 3 : 1;
            const int var = c < nshape ? 0 : c == c_pipe ? 1 : 2;
            const int all[5] = {ln, ln, ln, up, var};
            set_tiles(h, all);
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
            double warm[5], t[5];
            if ((rc = time_passes<R>(h, 1, warm)) || (rc = time_passes<R>(h, nsteps, t))) break;
            for (int p = 0; p < 5; ++p) tmin[c][p] = t[p] < tmin[c][p] ? t[p] : tmin[c][p];
        }
    if (!rc) {
        double best[5] = {1e300, 1e300, 1e300, 1e300, 1e300};
        for (int q = 0; q < nshape; ++q) {
            const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, q) : tile_lines_rt<float>(h->nz, q);
            const bool upd_ok = ((long long)h->nx * h->ny) % nl == 0;
            // a plan-1 candidate counts for a line pass only where its length
            // has a distinct radix-16 plan; the update takes shapes 0-2
            const bool use[4] = {q < 3 || plan_z, q < 3 || plan_y, q < 3 || plan_y, q < 3 && upd_ok};
            for (int p = 0; p < 4; ++p)
                if (use[p] && tmin[q][p] < best[p]) {
                    best[p] = tmin[q][p];
                    arg[p] = q;
                }
            best[4] = tmin[q][4] < best[4] ? tmin[q][4] : best[4];
        }
        if (c_pipe >= 0 && tmin[c_pipe][4] < best[4]) {
            best[4] = tmin[c_pipe][4];
            arg[4] = 1;
        }
        double five = best[0] + best[1] + best[2] + best[3] + best[4];
        // The kernels of a step share the last-level cache (a whole FP32
        // field of 128^3 x 12 fits in it), so a kernel's time depends on the
        // tiles of its neighbours, and the mix of per-kernel winners was
        // never timed as one step.  Contest: the mix against the two
        // candidates whose uniform choices had the smallest step sums, as
        // whole steps in two interleaved rounds; the smallest step wins.
        if (!rc) {
            int cand[3][5], nc = 0;
            for (int p = 0; p < 5; ++p) cand[nc][p] = arg[p];
            ++nc;
            int u1 = -1, u2 = -1;
            double s1 = 1e300, s2 = 1e300;
            for (int q = 0; q < nshape; ++q) {
                const double s = tmin[q][0] + tmin[q][1] + tmin[q][2] + tmin[q][3] + tmin[q][4];
                if (s < s1) { s2 = s1; u2 = u1; s1 = s; u1 = q; }
                else if (s < s2) { s2 = s; u2 = q; }
            }
            for (const int q : {u1, u2}) {
                if (q < 0) continue;
                const int nl = h->prec == PFDD_FP64 ? tile_lines_rt<double>(h->nz, q % 3)
                                                    : tile_lines_rt<float>(h->nz, q % 3);
                const bool upd_ok = ((long long)h->nx * h->ny) % nl == 0;
                const int u[5] = {q < 3 || plan_z ? q : q % 3, q < 3 || plan_y ? q : q % 3,
                                  q < 3 || plan_y ? q : q % 3, upd_ok ? q % 3 : arg[3], arg[4]};
                bool dup = false;
                for (int c = 0; c < nc; ++c) dup = dup || std::equal(u, u + 5, cand[c]);
                if (!dup) {
                    for (int p = 0; p < 5; ++p) cand[nc][p] = u[p];
                    ++nc;
                }
            }
            if (nc > 1) {
                double ts[3] = {1e300, 1e300, 1e300};
                for (int round = 0; round < 2 && !rc; ++round)
                    for (int c = 0; c < nc && !rc; ++c) {
                        set_tiles(h, cand[c]);
                        hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, h->stream, h->done, h->steps);
                        double warm[5], t[5];
                        if ((rc = time_passes<R>(h, 1, warm)) || (rc = time_passes<R>(h, nsteps, t))) break;
                        const double s = t[0] + t[1] + t[2] + t[3] + t[4];
                        ts[c] = s < ts[c] ? s : ts[c];
                    }
                int bc = 0;
                for (int c = 1; c < nc; ++c) bc = ts[c] < ts[bc] ? c : bc;
                if (!rc) {
                    for (int p = 0; p < 5; ++p) arg[p] = cand[bc][p];
                    five = ts[bc];
                }
            }
        }
        if (!rc && c_plane >= 0) {
            const double three = tmin[c_plane][0] + tmin[c_plane][3] + tmin[c_plane][4];
            if (three < five) arg[4] = 2;
        }
    }
    int rc2 = restore();
    if (!rc && !rc2) {
        set_tiles(h, arg);
        // the tuned launch shapes are final: capture pfdd_gl_step's graph now
        // (tol 0) rather than inside the caller's first long run, replay it
        // once from the restored field (which also keeps the GPU busy up to
        // the caller's first run instead of idling through the capture) and
        // restore again
        if (h->sg_on && !(rc = self_graph_ready(h, 0.0))) {
            const hipError_t e = hipGraphLaunch(h->sg[h->cur].exec, h->stream);
            if (e != hipSuccess) rc = fail(PFDD_ERR_HIP, "autotune graph replay: %s", hipGetErrorString(e));
            rc2 = restore();
        }
    }
    const hipError_t es = hipStreamSynchronize(h->stream);
    (void)hipFree(save);
    if (!rc) rc = rc2;
    if (!rc && es != hipSuccess) rc = fail(PFDD_ERR_HIP, "autotune: %s", hipGetErrorString(es));
    if (rc) {
        const int dflt[5] = {1, 1, 1, 1, 0};
        set_tiles(h, dflt);
        return rc;
    }
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

void self_graph_release(pfdd_gl *h)
{
    for (auto &g : h->sg) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
        g.graph = nullptr;
    }
}

void destroy_all(pfdd_gl *h)
{
    if (!h) return;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    self_graph_release(h);
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

// Iterations per replay of pfdd_gl_step's graph (even: the ping-pong buffer
// is the same before and after a replay).
constexpr int SG_ITERS = 10;

// Captures SG_ITERS GL iterations of h, starting from xi[par], on its own
// stream into h->sg[par].  Capturing enqueues nothing; h->cur and
// h->stepped are back where they started.
int self_graph_capture(pfdd_gl *h, double tol, int par)
{
    auto &g = h->sg[par];
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
    const int cur0 = h->cur;
    const bool stepped0 = h->stepped;
    h->cur = par;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = PFDD_OK;
    for (int it = 0; it < SG_ITERS && !rc; ++it)
        rc = h->prec == PFDD_FP64 ? step_once<double>(h, tol, nullptr) : step_once<float>(h, tol, nullptr);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    h->cur = cur0;
    h->stepped = stepped0;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(PFDD_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    g.graph = graph;
    HIPCHK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    // the graph of the other parity is first launched later, often inside a
    // timed run: its upload to the device is done here
    HIPCHK(hipGraphUpload(g.exec, h->stream));
    g.ver = h->version;
    g.tol = tol;
    return PFDD_OK;
}

// Makes h->sg[h->cur] current for tol.  Both parities are captured together:
// a run with an odd number of eager iterations (a short warm-up, an odd
// remainder) switches the parity, and capturing then inside the next long
// run cost about 0.45 ms on its first 20 iterations (128^3 x 12).
int self_graph_ready(pfdd_gl *h, double tol)
{
    auto fresh = [&](int par) { return h->sg[par].exec && h->sg[par].ver == h->version && h->sg[par].tol == tol; };
    if (fresh(h->cur)) return PFDD_OK;
    int rc = PFDD_OK;
    for (const int par : {h->cur, 1 - h->cur})
        if (!rc && !fresh(par)) rc = self_graph_capture(h, tol, par);
    return rc;
}

} // namespace

// ===================================================================== C ABI
extern "C" {

int pfdd_gl_version(void) { return (0 << 16) | 1; }

int pfdd_gl_stream_rates(int device, long long bytes, int reps, double *copy_GBps, double *add_GBps)
{
    if (bytes < 4096 || reps < 1 || !copy_GBps || !add_GBps) return fail(PFDD_ERR_ARG, "stream_rates: bad argument");
    if (hipSetDevice(device) != hipSuccess) return fail(PFDD_ERR_ARG, "stream_rates: no device %d", device);
    const long long n = bytes / 16;
    double2 *a = nullptr, *b = nullptr, *c = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PFDD_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && !rc) rc = fail(PFDD_ERR_HIP, "stream_rates %s: %s", what, hipGetErrorString(e));
    };
    chk(hipMalloc((void **)&a, n * 16), "alloc");
    if (!rc) chk(hipMalloc((void **)&b, n * 16), "alloc");
    if (!rc) chk(hipMalloc((void **)&c, n * 16), "alloc");
    if (!rc) chk(hipStreamCreate(&st), "stream");
    if (!rc) chk(hipEventCreate(&e0), "event");
    if (!rc) chk(hipEventCreate(&e1), "event");
    const unsigned nb = (unsigned)((n + 255) / 256);
    double rate[2] = {0.0, 0.0};
    for (int kind = 0; kind < 2 && !rc; ++kind) {
        auto launch = [&]() {
            if (kind == 0) hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, st, a, b, n);
            else hipLaunchKernelGGL(k_stream_add, dim3(nb), dim3(256), 0, st, a, c, b, n);
        };
        // first touch of the pages, then the clock ramp, then the timing
        for (int w = 0; w < 20; ++w) launch();
        chk(hipGetLastError(), "launch");
        chk(hipEventRecord(e0, st), "record");
        for (int r = 0; r < reps; ++r) launch();
        chk(hipEventRecord(e1, st), "record");
        chk(hipEventSynchronize(e1), "sync");
        float ms = 0.0f;
        if (!rc) chk(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (!rc && ms > 0) rate[kind] = (kind == 0 ? 2.0 : 3.0) * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (c) (void)hipFree(c);
    *copy_GBps = rate[0];
    *add_GBps = rate[1];
    return rc;
}

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
    {
        double *sp = nullptr; // the device's no-allocation scratch, resolved at set-up
        if ((rc = noalloc_scratch(d->device, &sp))) return bail(rc);
    }
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
    // the plane iteration: FP32 or FP64 grids with square y-z planes that fit
    // in LDS and an x pass of the plane layout for nx; not with stored B
    {
        const long long ncol = (long long)h->ny * h->nz, nplanes = (long long)h->SM * h->nx;
        const bool gen = d->core_model == PFDD_CORE_EXTENDED || d->core_model == PFDD_CORE_PYRII ||
                         d->core_model == PFDD_CORE_FOURIER;
        const bool f64 = h->prec == PFDD_FP64;
        const bool sup = f64 ? plane_supported<double>(h->ny, h->nz) : plane_supported<float>(h->ny, h->nz);
        h->nb3pl = f64 ? mid_pl_blocks<double>(h->nx, h->S, h->M, ncol) : mid_pl_blocks<float>(h->nx, h->S, h->M, ncol);
        h->plane_ok = sup && h->nb3pl > 0 && !h->stored_B;
        if (h->plane_ok) {
            h->nbA = f64 ? plane_blocks<double>(h->ny, h->nz, nplanes, h->ncu, 0)
                         : plane_blocks<float>(h->ny, h->nz, nplanes, h->ncu, 0);
            h->nbC = f64 ? plane_blocks<double>(h->ny, h->nz, nplanes, h->ncu, gen ? 2 : 1)
                         : plane_blocks<float>(h->ny, h->nz, nplanes, h->ncu, gen ? 2 : 1);
            const long long mx = h->nb3pl > nplanes ? h->nb3pl : nplanes;
            if (mx > h->pstride) h->pstride = mx;
        }
    }
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
    {
        const char *g = std::getenv("PFDD_GL_GRAPH");
        h->sg_on = !(g && g[0] == '0');
    }
    if (d->flags & PFDD_GL_FLAG_AUTOTUNE) {
        rc = h->prec == PFDD_FP64 ? autotune<double>(h, 4, nullptr) : autotune<float>(h, 4, nullptr);
        // (a cached choice returns without tuning: capture here; a no-op
        // when the tuner has just captured)
        if (!rc && h->sg_on) rc = self_graph_ready(h, 0.0);
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
    for (int p = 0; p < 3; ++p)
        if (choice[p] < 0 || choice[p] > 5)
            return fail(PFDD_ERR_ARG, "line-pass choice must be 0-5 (tile shape + 3 x radix plan)");
    if (choice[3] < 0 || choice[3] > 2) return fail(PFDD_ERR_ARG, "update tile shape choice must be 0, 1 or 2");
    if (choice[4] < 0 || choice[4] > 2) return fail(PFDD_ERR_ARG, "iteration choice must be 0, 1 or 2");
    if (choice[4] == 2 && !h->plane_ok)
        return fail(PFDD_ERR_UNSUPPORTED, "the plane iteration needs square y-z planes that fit in LDS "
                                          "(FP32 up to 128 x 128, FP64 up to 80 x 80), nx of 16..128 and no stored B");
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
    rc = h->prec == PFDD_FP64 ? autotune<double>(h, nsteps, choice) : autotune<float>(h, nsteps, choice);
    // a cached choice returns without tuning: its graph may still be stale
    return rc || !h->sg_on ? rc : self_graph_ready(h, 0.0);
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
    int eager = nsteps;
    // long runs replay a graph of SG_ITERS iterations (one launch instead of
    // 6 SG_ITERS), unless the caller is capturing this stream itself
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(h->stream, &cs));
    if (h->sg_on && nsteps >= 2 * SG_ITERS && cs == hipStreamCaptureStatusNone) {
        if ((rc = self_graph_ready(h, tol))) return rc;
        // SG_ITERS is even: every replay leaves the parity where it was
        for (int r = 0; r < nsteps / SG_ITERS; ++r) HIPCHK(hipGraphLaunch(h->sg[h->cur].exec, h->stream));
        h->stepped = true;
        eager = nsteps % SG_ITERS;
    }
    for (int s = 0; s < eager; ++s) {
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
    const int nb = XS_NB;
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

int pfdd_gl_xi_sums_noalloc(pfdd_gl *h, double *sum_re)
{
    int rc;
    if ((rc = check_handle(h)) || (rc = set_device(h))) return rc;
    if (!sum_re) return fail(PFDD_ERR_ARG, "null output");
    if (h->SM > PFDD_GL_NOALLOC_MAX_SYSTEMS)
        return fail(PFDD_ERR_UNSUPPORTED, "pfdd_gl_xi_sums_noalloc: %d systems exceed its scratch limit of %d",
                    h->SM, PFDD_GL_NOALLOC_MAX_SYSTEMS);
    double *part = nullptr;
    if ((rc = noalloc_scratch(h->device, &part))) return rc;
    double *sums = part + (long long)XS_NB * PFDD_GL_NOALLOC_MAX_SYSTEMS;
    if (h->prec == PFDD_FP64)
        hipLaunchKernelGGL((k_sys_sums<double>), dim3(XS_NB, h->SM), dim3(NTHR), 0, h->stream,
                           (const double2 *)h->xi[h->cur], h->N, part);
    else
        hipLaunchKernelGGL((k_sys_sums<float>), dim3(XS_NB, h->SM), dim3(NTHR), 0, h->stream,
                           (const float2 *)h->xi[h->cur], h->N, part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_sys_sums_final, dim3(h->SM), dim3(64), 0, h->stream, (const double *)part, XS_NB, sums);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sum_re, sums, sizeof(double) * h->SM, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
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
    // the plane iteration has no separate y passes (slots 1 and 3)
    const double k[5] = {2 * f, h->plane ? 0.0 : 2 * f, 2 * f + bsz, h->plane ? 0.0 : 2 * f, 3 * f};
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
    std::vector<hipStream_t> branch; // one capture stream per handle: parallel branches of the graph
    std::vector<hipEvent_t> join;    // fork / join events of the capture (one per handle, plus the fork)
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

// The graph forks at its root into one branch per handle (the handle's
// iterations, in order, on its own capture stream) and joins at the end, so
// the kernels of different handles -- e.g. the small grids of an ensemble --
// may run concurrently.
int graph_capture(pfdd_gl_graph *g, double tol)
{
    graph_release(g);
    const size_t n = g->hs.size();
    std::vector<hipStream_t> saved(n);
    for (size_t i = 0; i < n; ++i) {
        saved[i] = g->hs[i]->stream;
        g->hs[i]->stream = g->branch[i];
    }
    int rc = PFDD_OK;
    hipError_t e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        rc = fail(PFDD_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    } else {
        e = hipEventRecord(g->join[n], g->stream); // fork
        for (size_t i = 0; i < n && !rc && e == hipSuccess; ++i) {
            pfdd_gl *h = g->hs[i];
            e = hipStreamWaitEvent(h->stream, g->join[n], 0);
            for (int it = 0; it < g->iters && !rc && e == hipSuccess; ++it)
                rc = h->prec == PFDD_FP64 ? step_once<double>(h, tol, nullptr) : step_once<float>(h, tol, nullptr);
            if (!rc && e == hipSuccess) e = hipEventRecord(g->join[i], h->stream);
            if (!rc && e == hipSuccess) e = hipStreamWaitEvent(g->stream, g->join[i], 0);
        }
        if (!rc && e != hipSuccess) rc = fail(PFDD_ERR_HIP, "graph fork/join: %s", hipGetErrorString(e));
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
    g->branch.assign(n, nullptr);
    g->join.assign(n + 1, nullptr);
    hipError_t e = hipSetDevice(handles[0]->device);
    if (e == hipSuccess) e = hipStreamCreate(&g->stream);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = hipStreamCreate(&g->branch[i]);
    for (int i = 0; i <= n && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        pfdd_gl_graph_destroy(g);
        return fail(PFDD_ERR_HIP, "graph streams: %s", hipGetErrorString(e));
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
    for (hipStream_t b : g->branch)
        if (b) (void)hipStreamDestroy(b);
    for (hipEvent_t ev : g->join)
        if (ev) (void)hipEventDestroy(ev);
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
