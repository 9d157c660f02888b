// pfdd_mid.hip -- the x pass: FFT along x, Green's-function contraction,
// inverse FFT along x (stored-B and on-the-fly variants), and launchers.
#include "pfdd_kern.hpp"

namespace pfdd {

// Load the S lines of NP consecutive columns (T2[a][c0+q][0..N)) into the
// tile, lines ordered (a, q); column-contiguous runs of N elements.  Every
// load of a thread is issued before its first LDS write, in both precisions
// (for the 16-byte elements of the FP64 x pass at 128^3 x 12 this measured
// 191 -> 184 us against a strided loop of 4 loads in flight).
template <int N, int NL, int NP, int NTHR, typename XA, typename V>
__device__ __forceinline__ void tile_load_cols(const Tile<V> lds, const V *__restrict__ T2, long long c0, long long ncol,
                                               long long ntot, int tid)
{
    constexpr int TOT = NL * N, PER = (TOT + NTHR - 1) / NTHR;
    {
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
            if (TOT % NTHR == 0 || e < TOT) lds.st(XA::at(e % N, e / N), r[t]);
        }
    }
}

// --------------------------------------------- pass 3: x FFT + contraction
// Columns per x-pass tile: the largest power of two whose S*NP lines of N
// elements stay within `target` elements.  FP64 with 12 systems takes twice
// the tile of the other FP64 cases: at 128^3 x 12 two columns per block (one
// thread per k-point, no duplicated contraction work, 3 blocks per CU) ran
// the pass 12 % faster than one column (two threads per k-point, 5 blocks).
// So do 512-point FP64 lines (two columns of 3 systems instead of one:
// 512 x 512 x 256 x 3, 2.52 -> 2.27 ms; in FP32 four columns instead of two
// were slower).
template <typename R, int N, int S> constexpr int mid_pairs()
{
    constexpr int target = sizeof(R) == 8 ? ((S == 12 || N >= 512) ? 4096 : 2048) : 4096;
    int np = 1;
    while (np * 2 * S * N <= target) np *= 2;
    // the padded tile stays within 60 KB of LDS (the constants and the
    // reduction scratch of the pass come on top)
    while (np > 1 && (long long)N * (S * np + 1) * 2 * (int)sizeof(R) > 61440) np /= 2;
    return np;
}

// Tile addressing of the on-the-fly x pass: Swz8 for its 24-line FP64 tiles
// (12 systems x 2 columns at 96 and 128 points); unpadded rows for FP64
// lines of 512 points, whose padded tile (57.6 KB) leaves room for only two
// blocks per CU; RowPad otherwise.
template <typename R, int N, int S> struct XLay {
    static constexpr int NL = S * mid_pairs<R, N, S>();
    typedef typename std::conditional<
        (sizeof(R) == 8 && NL == 24 && N % 16 == 0), Swz8<NL>,
        typename std::conditional<(sizeof(R) == 8 && N >= 512), RowPad<NL, 0>, RowPad<NL>>::type>::type type;
};

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
    double el = 0.0; // few k-points per thread; the block and grid sums are compensated
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
    block_sum_comp<1, NTHR>(v, red, tid);
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

// Hands ghat_a for a = H*SA .. H*SA+SA-1 of one k-point to sink(aa, re, im,
// xihat_a), computed in T (double in the FP64 path, float in the FP32 mode).
// xihat_a is the value read for the contraction, kept in registers so that
// the sink does not read the tile again (FP64 x pass at 128^3 x 12: 189 ->
// 186 us).
template <typename T, int S, int SA, int H, int NP, typename XA, typename V, typename Sink>
__device__ __forceinline__ void otf_point(const Tile<V> lds, int p, int q, const T f[3], bool nzf, T c2, T c4,
                                          const T *__restrict__ cst, const GConst<T> &G, Sink sink)
{
    // compiler-only barriers keep the (LDS-broadcast) constants from being
    // hoisted and held in registers across points and systems
    __asm__ volatile("" ::: "memory");
    T Pr[6] = {}, Pi[6] = {}; // Sig, then P
    V xk[SA];
#pragma unroll
    for (int b = 0; b < S; ++b) {
        const V v = lds.ld(XA::at(p, b * NP + q));
        if (b >= H * SA && b < H * SA + SA) xk[b - H * SA] = v;
        const T xr = (T)v.x, xi = (T)v.y;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            Pr[c] += cst[b * 12 + c] * xr;
            Pi[c] += cst[b * 12 + c] * xi;
        }
    }
    if (nzf) {
        // symmetric components xx, yy, zz, xy, xz, yz
        const T wr[3] = {Pr[0] * f[0] + Pr[3] * f[1] + Pr[4] * f[2], Pr[3] * f[0] + Pr[1] * f[1] + Pr[5] * f[2],
                              Pr[4] * f[0] + Pr[5] * f[1] + Pr[2] * f[2]};
        const T wi[3] = {Pi[0] * f[0] + Pi[3] * f[1] + Pi[4] * f[2], Pi[3] * f[0] + Pi[1] * f[1] + Pi[5] * f[2],
                              Pi[4] * f[0] + Pi[5] * f[1] + Pi[2] * f[2]};
        const T Vr = f[0] * wr[0] + f[1] * wr[1] + f[2] * wr[2];
        const T Vi = f[0] * wi[0] + f[1] * wi[1] + f[2] * wi[2];
        // c2 multiplies (w_i f_j + w_j f_i), c4 multiplies V f_i f_j
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
        if (!nzf) gr = gi = (T)0;
        sink(aa, gr, gi, xk[aa]);
    }
}

// f = four^T k~ of k-point (p, col) (frec, fft_fftw_slab.cpp:280-367) and the
// scalars of the Green's function G = (2 I/f^2 - f f/((1-nu) f^4))/(2 C44)
// (Bmatrix, :577-634), c2 = 1/(2 C44 f^2) and c4 = 1/((1-nu) 2 C44 f^4), all
// formed in FP64 and rounded once to T: in the FP32 mode B(k) is built in
// FP64 and only the contraction with xihat runs in FP32.  nzf: f != 0 (B(0)
// = 0).  Column col is (k, j) = (col / ny, col % ny) in the line layout of the
// five-kernel iteration and (j, k) = (col / nz, col % nz) in the plane layout.
template <int N, int NP, int S, typename T, bool PL = false>
__device__ __forceinline__ void otf_freq(const GParams &G, long long col, int p, T f[3], bool &nzf, T &c2, T &c4)
{
    const int c = (int)col; // column of one member: ny * nz < 2^31
    int j, k;
    if constexpr (PL) {
        j = c / G.nz;
        k = c - j * G.nz;
    } else {
        k = c / G.ny;
        j = c - k * G.ny;
    }
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
    const double f2 = fd[0] * fd[0] + fd[1] * fd[1] + fd[2] * fd[2];
    nzf = f2 > 0;
    const double if2 = nzf ? 1.0 / f2 : 0.0;
    c2 = (T)(if2 * G.inv2C44);
    c4 = (T)(G.inv1mnu * if2 * if2 * G.inv2C44);
}

// Minimum waves per SIMD asked of the x pass (one wave per SIMD per resident
// block of 256 threads).  The 12-system FP64 cases ask for 5, set when their
// tiles were one column wide; with two-column tiles their LDS allows 3 blocks
// per CU, which the compiler then reaches (measured, and kept: the register
// allocation the hint produces is the one timed).  The 3-system passes ask
// for 3 in FP64 (168 instead of 172-190 VGPRs, a few spilled; their tiles
// allow 3-8 blocks): 3-system batches at 48^3 and 64^3 ran their x pass
// 80 -> 67 and 193 -> 179 us, and 512-point lines (unpadded tile, 49 KB)
// 2.25 -> 1.81 ms.  In FP32 only the 512-point lines ask for 5 (1.17 -> 1.11
// ms); at 48-96 points 5 or 6 waves (96 or 80 VGPRs, with spills) were
// slower than the compiler's 4.  Other system counts are not constrained.
template <typename R, int N, int S> constexpr int otf_waves()
{
    if (S == 12 && sizeof(R) == 8 && N >= 32 && N <= 128) return 5;
    if (S == 3) return sizeof(R) == 8 ? 3 : (N >= 512 ? 5 : 1);
    return 1;
}

// The transform of one x-pass tile held in LDS (NP columns x S systems):
// forward FFT along x, on-the-fly contraction ghat_a = sum_b B_ab xihat_b,
// inverse FFT.  Returns this thread's share of sum Re(conj(xihat_a) ghat_a).
template <typename R, int N, int S, bool PL = false>
__device__ __forceinline__ double otf_tile(const Tile<typename cvec<R>::type> lds,
                                           const typename cvec<R>::type *__restrict__ tw, const GParams &G,
                                           const GConst<R> &gc, const R *cst, long long c0, long long ncol, int tid)
{
    typedef typename cvec<R>::type V;
    typedef R T;
    constexpr int NP = mid_pairs<R, N, S>();
    constexpr int NL = S * NP;
    typedef typename XLay<R, N, S>::type XA;
    constexpr int TPP = otf_tpp<R, N, S>();
    constexpr int SA = S / TPP;
    constexpr int NPT = NP * N; // k-points per tile
    fft_stages_x<N, 1, NL, NTHR, false, XA>(lds, tw, tid);
    // the elastic-energy term Re(conj(xihat_a) ghat_a) is taken as each ghat_a
    // replaces xihat_a in the tile
    double el = 0.0; // a thread handles one or two k-points; the block and grid sums are compensated
    if constexpr (TPP == 1) {
        // one thread per k-point: it reads all xihat_b of the point, then
        // overwrites them with ghat_a
        for (int pt = tid; pt < NPT; pt += NTHR) {
            const int q = pt / N, p = pt % N;
            if (c0 + q >= ncol) continue;
            T f[3], c2, c4;
            bool nzf;
            otf_freq<N, NP, S, T, PL>(G, c0 + q, p, f, nzf, c2, c4);
            otf_point<T, S, SA, 0, NP, XA>(lds, p, q, f, nzf, c2, c4, cst, gc, [&](int a, T gr, T gi, V x) {
                const int idx = XA::at(p, a * NP + q);
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
        T f[3], c2, c4;
        bool nzf;
        otf_freq<N, NP, S, T, PL>(G, c0 + q, p, f, nzf, c2, c4);
        T g[SA][2];
        auto keep = [&](int aa, T gr, T gi, V) {
            g[aa][0] = gr;
            g[aa][1] = gi;
        };
        if (ok) {
            if (__builtin_amdgcn_readfirstlane(h) == 0)
                otf_point<T, S, SA, 0, NP, XA>(lds, p, q, f, nzf, c2, c4, cst, gc, keep);
            else
                otf_point<T, S, SA, 1, NP, XA>(lds, p, q, f, nzf, c2, c4, cst, gc, keep);
        }
        __syncthreads();
        if (ok) {
#pragma unroll
            for (int aa = 0; aa < SA; ++aa) {
                const int idx = XA::at(p, (h * SA + aa) * NP + q);
                const V x = lds.ld(idx);
                el += (double)x.x * g[aa][0] + (double)x.y * g[aa][1];
                lds.st(idx, mk<V>((R)g[aa][0], (R)g[aa][1]));
            }
        }
    }
    __syncthreads();
    fft_stages_x<N, 1, NL, NTHR, true, XA>(lds, tw, tid);
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
    constexpr int NL = S * NP, TOT = NL * N;
    typedef typename XLay<R, N, S>::type XA;
    PFDD_TILE(V, lds, XA::size(N));
    __shared__ R cst[S * 12];
    __shared__ double red[NTHR / 64];
    const int tid = threadIdx.x;
    // member of a batch (bpm blocks per member), first column of the block
    const int mem = blockIdx.x / bpm;
    const long long c0 = (long long)(blockIdx.x - mem * bpm) * NP;
    T2 += (long long)mem * S * ntot;
    GConst<R> gc;
    otf_consts<R, S>(cst, gc, G, tid);
    tile_load_cols<N, NL, NP, NTHR, XA>(lds, T2, c0, ncol, ntot, tid);
    __syncthreads();
    double v[1] = {otf_tile<R, N, S>(lds, tw, G, gc, cst, c0, ncol, tid)};
#pragma unroll 4
    for (int e = tid; e < TOT; e += NTHR) {
        const int l = e / N, p = e % N;
        const int a = l / NP, q = l % NP;
        if (c0 + q < ncol) st_stream(&T2[(long long)a * ntot + (c0 + q) * N + p], lds.ld(XA::at(p, l)));
    }
    block_sum_comp<1, NTHR>(v, red, tid);
    if (tid == 0) part_el[blockIdx.x] = v[0];
}

// x pass over the plane layout T[a][i][j][k] of the two-pass iteration
// (pfdd_plane.hip).  Column c = j * nz + k is strided by ny * nz along i and
// the NP columns of a tile are consecutive in k, so each (a, i) of a tile is
// a run of NP elements, often shorter than a 128-byte line.  The other tiles
// of the line are put on the same XCD at about the same time: blocks are
// dispatched to the 8 XCDs round-robin (block b to XCD b % 8), and block b
// takes tile 64 (b / 64) + 8 (b % 8) + (b / 8) % 8, so the 8 consecutive tiles
// 8 r .. 8 r + 7 of a group of 64 run on one XCD and share each line through
// its L2.  The grid is a multiple of 64 blocks; blocks past the last tile
// contribute a zero partial.
template <typename R, int N, int S>
__global__ void __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(otf_waves<R, N, S>(), 8)))
    k_mid_otf_pl(typename cvec<R>::type *__restrict__ T, GParams G, long long ncol, long long ntot, int bpm,
                 long long ntiles, const typename cvec<R>::type *__restrict__ tw, double *__restrict__ part_el,
                 const int *__restrict__ done)
{
    typedef typename cvec<R>::type V;
    if (*done) return;
    constexpr int NP = mid_pairs<R, N, S>();
    constexpr int NL = S * NP, TOT = NL * N, PER = (TOT + NTHR - 1) / NTHR;
    typedef typename XLay<R, N, S>::type XA;
    PFDD_TILE(V, lds, XA::size(N));
    __shared__ R cst[S * 12];
    __shared__ double red[NTHR / 64];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const long long t = ((long long)(b >> 6) << 6) + ((b & 7) << 3) + ((b >> 3) & 7);
    if (t >= ntiles) {
        if (tid == 0) part_el[b] = 0.0;
        return;
    }
    const int mem = (int)(t / bpm);
    const long long c0 = (t - (long long)mem * bpm) * NP;
    T += (long long)mem * S * ntot;
    GConst<R> gc;
    otf_consts<R, S>(cst, gc, G, tid);
    // element e of the tile: column q fastest, then position p (= i), then
    // system a; every load of a thread is in flight before its LDS writes.
    // Plain (temporal) accesses: the lines are shared with the other tiles.
    V r[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = tid + k * NTHR;
        const int q = e % NP, p = (e / NP) % N, a = e / (NP * N);
        if ((TOT % NTHR == 0 || e < TOT) && c0 + q < ncol) r[k] = T[(long long)a * ntot + (long long)p * ncol + c0 + q];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = tid + k * NTHR;
        const int q = e % NP, p = (e / NP) % N, a = e / (NP * N);
        if (TOT % NTHR == 0 || e < TOT) lds.st(XA::at(p, a * NP + q), r[k]);
    }
    __syncthreads();
    double v[1] = {otf_tile<R, N, S, true>(lds, tw, G, gc, cst, c0, ncol, tid)};
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = tid + k * NTHR;
        const int q = e % NP, p = (e / NP) % N, a = e / (NP * N);
        if ((TOT % NTHR == 0 || e < TOT) && c0 + q < ncol)
            T[(long long)a * ntot + (long long)p * ncol + c0 + q] = lds.ld(XA::at(p, a * NP + q));
    }
    block_sum_comp<1, NTHR>(v, red, tid);
    if (tid == 0) part_el[b] = v[0];
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
    constexpr int NL = S * NP, TOT = NL * N, PER = (TOT + NTHR - 1) / NTHR;
    typedef typename XLay<R, N, S>::type XA;
    PFDD_TILE(V, lds, XA::size(N));
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
            if (TOT % NTHR == 0 || e < TOT) lds.st(XA::at(e % N, e / N), r[k]);
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
            if (c0 + q < ncol) T2m[(long long)a * ntot + (c0 + q) * N + p] = lds.ld(XA::at(p, l));
        }
        __syncthreads();
    }
    double v[1] = {el};
    block_sum_comp<1, NTHR>(v, red, threadIdx.x);
    if (threadIdx.x == 0) part_el[blockIdx.x] = v[0];
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


// x pass of the plane layout: the grid sizes of the plane iteration.
#define PFDD_PL_SIZES(X) X(16) X(32) X(48) X(64) X(80) X(96) X(128)

template <typename R, int S>
hipError_t launch_mid_pl_s(int n, int M, void *T, const GParams &G, long long ncol, long long ntot, const void *tw,
                           double *part_el, const int *done, hipStream_t st, int *nblocks)
{
    typedef typename cvec<R>::type V;
#define PFDD_CASE(n_)                                                                                  \
    if (n == n_) {                                                                                     \
        if constexpr (mid_fits<R, n_, S>()) {                                                          \
            constexpr int NP = mid_pairs<R, n_, S>();                                                  \
            const int bpm = (int)((ncol + NP - 1) / NP);                                               \
            const long long nt = (long long)M * bpm;                                                   \
            *nblocks = (int)((nt + 63) / 64 * 64);                                                     \
            PFDD_LAUNCH((k_mid_otf_pl<R, n_, S>), dim3((unsigned)*nblocks), dim3(NTHR), st, (V *)T, G,  \
                        ncol, ntot, bpm, nt, (const V *)tw, part_el, done);                            \
            return hipGetLastError();                                                                  \
        }                                                                                              \
        return hipErrorInvalidValue;                                                                   \
    }
    PFDD_PL_SIZES(PFDD_CASE)
#undef PFDD_CASE
    return hipErrorInvalidValue;
}

template <typename R>
hipError_t launch_mid_pl(int n, int S, int M, void *T, const GParams &G, long long ncol, long long ntot,
                         const void *tw, double *part_el, const int *done, hipStream_t st, int *nblocks)
{
    switch (S) {
    case 1: return launch_mid_pl_s<R, 1>(n, M, T, G, ncol, ntot, tw, part_el, done, st, nblocks);
    case 2: return launch_mid_pl_s<R, 2>(n, M, T, G, ncol, ntot, tw, part_el, done, st, nblocks);
    case 3: return launch_mid_pl_s<R, 3>(n, M, T, G, ncol, ntot, tw, part_el, done, st, nblocks);
    case 4: return launch_mid_pl_s<R, 4>(n, M, T, G, ncol, ntot, tw, part_el, done, st, nblocks);
    case 6: return launch_mid_pl_s<R, 6>(n, M, T, G, ncol, ntot, tw, part_el, done, st, nblocks);
    case 12: return launch_mid_pl_s<R, 12>(n, M, T, G, ncol, ntot, tw, part_el, done, st, nblocks);
    }
    return hipErrorInvalidValue;
}

// Blocks of the plane-layout x pass (a multiple of 64), 0 if unsupported.
template <typename R>
int mid_pl_blocks(int n, int S, int M, long long ncol)
{
    bool ok = false;
#define PFDD_CASE(n_) if (n == n_) ok = true;
    PFDD_PL_SIZES(PFDD_CASE)
#undef PFDD_CASE
    if (!ok || !mid_supported<R>(n, S)) return 0;
    const long long nt = (long long)M * mid_blocks<R>(n, S, ncol);
    return (int)((nt + 63) / 64 * 64);
}

#define PFDD_INST(R)                                                                                              \
    template int mid_blocks<R>(int, int, long long);                                                              \
    template bool mid_supported<R>(int, int);                                                                     \
    template hipError_t launch_mid<R>(int, int, int, void *, const void *, long long, long long, const void *,     \
                                      double *, const int *, hipStream_t);                                        \
    template hipError_t launch_mid_otf<R>(int, int, int, void *, const GParams &, long long, long long,            \
                                          const void *, double *, const int *, hipStream_t, int, int, int *);      \
    template hipError_t launch_mid_pl<R>(int, int, int, void *, const GParams &, long long, long long,             \
                                         const void *, double *, const int *, hipStream_t, int *);                  \
    template int mid_pl_blocks<R>(int, int, int, long long);
PFDD_INST(double)
PFDD_INST(float)
#undef PFDD_INST

} // namespace pfdd
