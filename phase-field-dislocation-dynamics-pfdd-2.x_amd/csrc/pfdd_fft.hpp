// pfdd_fft.hpp -- device building blocks of the batched 3-D FFT used by the GL
// step: complex arithmetic, radix-2/3/4/8 butterflies, mixed-radix Stockham
// stages over a tile of lines held in LDS, and coalesced tile load/store.
//
// The transform is FFTW 2's (fft_fftw_slab.cpp:1962,2021): unnormalised,
// forward sign -1, backward sign +1.
//
// Tile layout in LDS: element (line l, position p) at index p * (NL + 1) + l.
// Butterflies are mapped line-fastest across threads, so a wave touches
// consecutive LDS addresses; the +1 pad keeps position-fastest accesses (the
// coalesced load/store of contiguous lines) free of bank conflicts.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfdd {

template <typename R> struct cvec;
template <> struct cvec<double> { typedef double2 type; };
template <> struct cvec<float> { typedef float2 type; };

template <typename V> __device__ __forceinline__ V mk(decltype(V{}.x) x, decltype(V{}.x) y)
{
    V r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
// a * w (forward) or a * conj(w) (inverse)
template <bool INV, typename V> __device__ __forceinline__ V cmulw(V a, V w)
{
    const auto wy = INV ? -w.y : w.y;
    return mk<V>(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
}
// a * e^{-i pi/2} = -i a (forward), a * e^{+i pi/2} = +i a (inverse)
template <bool INV, typename V> __device__ __forceinline__ V cmul_mi(V a)
{
    return INV ? mk<V>(-a.y, a.x) : mk<V>(a.y, -a.x);
}


// ------------------------------------------------------------ LDS tiles
// Storage of a tile of complex values in LDS.  float2 values are kept
// interleaved (8-byte accesses); double2 values are kept as two planes of
// doubles (re, im) so that every LDS access is 8 bytes wide: 16-byte
// accesses to the same tile patterns measured heavy bank conflicts.
// Streaming accesses of the field passes.  Every element is read once and
// written once per pass, so the loads carry the nontemporal hint (measured on
// MI355X at 128^3 x 12: FP32 step -17 %, FP64 -6 %).  Stores carry it for
// 16-byte elements only: for 8-byte elements it slowed the FP32 line passes
// by a quarter, for 16-byte elements it gained another 3 %.
template <typename V> __device__ __forceinline__ V ld_stream(const V *p)
{
    typedef decltype(V{}.x) E;
    typedef E E2 __attribute__((ext_vector_type(2)));
    const E2 t = __builtin_nontemporal_load(reinterpret_cast<const E2 *>(p));
    return mk<V>(t.x, t.y);
}
template <typename V> __device__ __forceinline__ void st_stream(V *p, V v)
{
    if constexpr (sizeof(V) >= 16) {
        typedef decltype(V{}.x) E;
        typedef E E2 __attribute__((ext_vector_type(2)));
        E2 t;
        t.x = v.x;
        t.y = v.y;
        __builtin_nontemporal_store(t, reinterpret_cast<E2 *>(p));
    } else {
        *p = v;
    }
}

template <typename V> struct Tile;
template <> struct Tile<float2> {
    float2 *p;
    __device__ __forceinline__ float2 ld(int i) const { return p[i]; }
    __device__ __forceinline__ void st(int i, float2 v) const { p[i] = v; }
};
template <> struct Tile<double2> {
    double *re, *im;
    __device__ __forceinline__ double2 ld(int i) const { return make_double2(re[i], im[i]); }
    __device__ __forceinline__ void st(int i, double2 v) const
    {
        re[i] = v.x;
        im[i] = v.y;
    }
};
__device__ __forceinline__ Tile<float2> make_tile(float *mem, int) { return Tile<float2>{(float2 *)mem}; }
__device__ __forceinline__ Tile<double2> make_tile(double *mem, int count)
{
    return Tile<double2>{mem, mem + count};
}
// Declares `name` (a Tile<V>) over COUNT complex values of static LDS.
#define PFDD_TILE(V, name, COUNT)                                                \
    __shared__ __align__(16) decltype(V{}.x) name##_mem[2 * (COUNT)];              \
    const Tile<V> name = make_tile(name##_mem, (COUNT))

// ---------------------------------------------------------------- butterflies
template <bool INV, typename V> __device__ __forceinline__ void dft2(V *v)
{
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV, typename V> __device__ __forceinline__ void dft4(V *v)
{
    V a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    V a2 = cadd(v[1], v[3]), a3 = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
}

template <bool INV, typename V> __device__ __forceinline__ void dft8(V *v)
{
    typedef decltype(V{}.x) R;
    V e[4] = {v[0], v[2], v[4], v[6]};
    V o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    const R s = (R)0.707106781186547524400844362104849039;
    V t1, t3;
    if (INV) {
        t1 = mk<V>(s * (o[1].x - o[1].y), s * (o[1].x + o[1].y));
        t3 = mk<V>(-s * (o[3].x + o[3].y), s * (o[3].x - o[3].y));
    } else {
        t1 = mk<V>(s * (o[1].x + o[1].y), s * (o[1].y - o[1].x));
        t3 = mk<V>(s * (o[3].y - o[3].x), -s * (o[3].x + o[3].y));
    }
    V t2 = cmul_mi<INV>(o[2]);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
}

template <bool INV, typename V> __device__ __forceinline__ void dft3(V *v)
{
    typedef decltype(V{}.x) R;
    const R h = (R)0.866025403784438646763723170752936183; // sqrt(3)/2
    V t = cadd(v[1], v[2]);
    V d = csub(v[1], v[2]);
    V y0 = cadd(v[0], t);
    V m = mk<V>(v[0].x - (R)0.5 * t.x, v[0].y - (R)0.5 * t.y);
    // forward: y1 = m - i h d, y2 = m + i h d ; inverse swaps
    V ihd = mk<V>(-h * d.y, h * d.x); // +i h d
    v[0] = y0;
    if (INV) {
        v[1] = cadd(m, ihd);
        v[2] = csub(m, ihd);
    } else {
        v[1] = csub(m, ihd);
        v[2] = cadd(m, ihd);
    }
}

// Length-5 DFT from the symmetric pairs t = x1 + x4, x2 + x3 and the
// antisymmetric ones x1 - x4, x2 - x3 (c_k = cos(2 pi k/5), s_k = sin(2 pi k/5)):
//   X0 = x0 + t1 + t2,  a1 = x0 + c1 t1 + c2 t2,  a2 = x0 + c2 t1 + c1 t2,
//   b1 = s1 d1 + s2 d2, b2 = s2 d1 - s1 d2,
//   forward: X1 = a1 - i b1, X4 = a1 + i b1, X2 = a2 - i b2, X3 = a2 + i b2;
//   inverse: the signs of i swap.
template <bool INV, typename V> __device__ __forceinline__ void dft5(V *v)
{
    typedef decltype(V{}.x) R;
    const R c1 = (R)0.309016994374947424102293417182819059;  // cos(2 pi/5)
    const R c2 = (R)-0.809016994374947424102293417182819059; // cos(4 pi/5)
    const R s1 = (R)0.951056516295153572116439333379382143;  // sin(2 pi/5)
    const R s2 = (R)0.587785252292473129168705954639072769;  // sin(4 pi/5)
    const V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const V d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    const V x0 = v[0];
    const V a1 = mk<V>(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
    const V a2 = mk<V>(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
    const V b1 = mk<V>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
    const V b2 = mk<V>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
    const V ib1 = mk<V>(-b1.y, b1.x), ib2 = mk<V>(-b2.y, b2.x); // i b
    v[0] = cadd(x0, cadd(t1, t2));
    if (INV) {
        v[1] = cadd(a1, ib1);
        v[4] = csub(a1, ib1);
        v[2] = cadd(a2, ib2);
        v[3] = csub(a2, ib2);
    } else {
        v[1] = csub(a1, ib1);
        v[4] = cadd(a1, ib1);
        v[2] = csub(a2, ib2);
        v[3] = cadd(a2, ib2);
    }
}

// y * (c - i s) forward, y * (c + i s) inverse
template <bool INV, typename V> __device__ __forceinline__ V crot(V y, decltype(V{}.x) c, decltype(V{}.x) s)
{
    if (INV) return mk<V>(y.x * c - y.y * s, y.y * c + y.x * s);
    return mk<V>(y.x * c + y.y * s, y.y * c - y.x * s);
}

// Length-16 DFT as 4 x 4 (n = 4 n1 + n2, k = k1 + 4 k2): length-4 DFTs over
// n1, the twiddles w16^(n2 k1) with w16^m = cos(2 pi m/16) -+ i sin(2 pi m/16),
// length-4 DFTs over n2.
template <bool INV, typename V> __device__ __forceinline__ void dft16(V *v)
{
    typedef decltype(V{}.x) R;
    const R c1 = (R)0.923879532511286756128183189396788933; // cos(pi/8)
    const R s1 = (R)0.382683432365089771728459984030398867; // sin(pi/8)
    const R r = (R)0.707106781186547524400844362104849039;  // cos(pi/4)
    V y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        V t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft4<INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
    y[1][1] = crot<INV>(y[1][1], c1, s1);   // m = 1
    y[1][2] = crot<INV>(y[1][2], r, r);     // m = 2
    y[1][3] = crot<INV>(y[1][3], s1, c1);   // m = 3
    y[2][1] = crot<INV>(y[2][1], r, r);     // m = 2
    y[2][2] = cmul_mi<INV>(y[2][2]);        // m = 4
    y[2][3] = crot<INV>(y[2][3], -r, r);    // m = 6
    y[3][1] = crot<INV>(y[3][1], s1, c1);   // m = 3
    y[3][2] = crot<INV>(y[3][2], -r, r);    // m = 6
    y[3][3] = crot<INV>(y[3][3], -c1, -s1); // m = 9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        V t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4<INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

template <int RAD, bool INV, typename V> __device__ __forceinline__ void dft(V *v)
{
    if constexpr (RAD == 2) dft2<INV>(v);
    else if constexpr (RAD == 3) dft3<INV>(v);
    else if constexpr (RAD == 4) dft4<INV>(v);
    else if constexpr (RAD == 5) dft5<INV>(v);
    else if constexpr (RAD == 8) dft8<INV>(v);
    else if constexpr (RAD == 16) dft16<INV>(v);
    else static_assert(RAD == 2, "unsupported radix");
}

constexpr int pick_radix(int m)
{
    return m % 8 == 0 ? 8 : m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : m % 5 == 0 ? 5 : m;
}

// Lines per LDS tile of a line pass: the largest power of two whose tile
// holds at most (1024 << TQ) FP64 or (2048 << TQ) FP32 elements, at least one
// line, and whose padded tile fits 60 KiB of LDS (TQ 0, 1) or 80 KiB (TQ 2:
// two blocks per CU of the 160 KiB).  TQ is the tile-shape choice of the
// autotuner (0, 1, 2).  The larger cap matters for long lines, whose strided
// passes move runs of NL elements: 512-point FP64 lines get 8 lines (128-byte
// runs) instead of 4.
template <typename R, int N, int TQ, long long CAP = (TQ == 2 ? 81920 : 61440)> constexpr int tile_lines_q()
{
    constexpr int target = (sizeof(R) == 8 ? 1024 : 2048) << TQ;
    int nl = 1;
    while (nl * 2 * N <= target) nl *= 2;
    while (nl > 1 && (long long)N * (nl + 1) * 2 * (int)sizeof(R) > CAP) nl /= 2;
    return nl;
}

// Lines per tile of the update kernels (inverse z pass + update, and its
// adjoint): the same shapes with the 60 KiB cap for all three.  They hold
// per-thread state for every cell of the tile they update (old values, sums),
// and 80 KiB tiles (up to 32 cells per thread) measured 30-40 % slower.
template <typename R, int N, int TQ> constexpr int tile_lines_u() { return tile_lines_q<R, N, TQ, 61440>(); }

template <typename R, int N> constexpr int tile_lines() { return tile_lines_q<R, N, 1>(); }

// Addressing of a tile of NL lines: element (position p, line l) at
// XA::at(p, l), XA::size(n) elements for lines of n.  The Stockham stages
// address through rd<NBL>(j, r, l) = at(j + r NBL, l) and wr<NS, RAD>(j, r, l)
// = at((j / NS) NS RAD + j % NS + r NS, l), r a compile-time index, so that a
// layout can fold the r-dependent part into the instruction's offset.
// RowPad: rows of NL + PAD elements; a pad of 1 puts position-fastest
// accesses (stride NL + 1, odd) on distinct banks, PAD 0 saves the LDS.
template <int NL, int PAD = 1> struct RowPad {
    static constexpr int size(int n) { return n * (NL + PAD); }
    __device__ __forceinline__ static int at(int p, int l) { return p * (NL + PAD) + l; }
    template <int NBL> __device__ __forceinline__ static int rd(int j, int r, int l) { return at(j + r * NBL, l); }
    template <int NS, int RAD> __device__ __forceinline__ static int wr(int j, int r, int l)
    {
        return at((j / NS) * NS * RAD + j % NS + r * NS, l);
    }
};

// Swz8: rows of NL (a multiple of 8) elements without the pad, 8 elements
// of gap after every 16 rows, and the low three bits of the line index
// XORed with (p >> 1) & 7.  For 8-byte accesses served 16 lanes at a time
// from 32 four-byte banks, a tile of NL = 24 lines then has no bank
// conflicts in any of its access orders: line-fastest runs (the FFT stages)
// are contiguous within each row and the gap separates rows 16 apart (the
// radix-16 stage writes row 16 j + r next to row 16 (j + 1) + r);
// position-fastest runs of 16 (loads, stores, the contraction: stride NL)
// are spread by the XOR.  RowPad's 24-line rows conflict on the 16-lane
// groups that straddle two rows (one extra cycle per group).  The XOR
// depends on p mod 16 only, so row offsets that are multiples of 16 fold
// into constants.
template <int NL> struct Swz8 {
    static_assert(NL % 8 == 0, "Swz8 permutes groups of 8 lines");
    static constexpr int ROW16 = 16 * NL + 8; // 16 rows and their gap
    static constexpr int size(int n) { return n * NL + ((n + 15) / 16) * 8; }
    __device__ __forceinline__ static int at(int p, int l) { return p * NL + (p >> 4) * 8 + (l ^ ((p >> 1) & 7)); }
    template <int NBL> __device__ __forceinline__ static int rd(int j, int r, int l)
    {
        if constexpr (NBL % 16 == 0) {
            return at(j, l) + (r * NBL / 16) * ROW16;
        } else if constexpr (NBL == 8) {
            // j < 8: row j + 8 r is row j or j + 8 of block r / 2
            return at(j + 8 * (r & 1), l) + (r >> 1) * ROW16;
        } else {
            return at(j + r * NBL, l);
        }
    }
    template <int NS, int RAD> __device__ __forceinline__ static int wr(int j, int r, int l)
    {
        if constexpr (NS % 16 == 0) {
            return at((j / NS) * NS * RAD + j % NS, l) + (r * NS / 16) * ROW16;
        } else if constexpr (NS == 1 && RAD == 16) {
            // row 16 j + r: block j, row r within it
            return j * ROW16 + r * NL + (l ^ ((r >> 1) & 7));
        } else {
            return at((j / NS) * NS * RAD + j % NS + r * NS, l);
        }
    }
};

// One Stockham (auto-sort) radix-RAD stage over all NL lines of the tile,
// in place in LDS: every thread reads all its inputs, barrier, writes.
//   input  (line l, j + r*N/RAD), r < RAD, for butterfly j < N/RAD
//   twiddle w^{r*(j mod NS)}, w = exp(-+2 pi i / (NS*RAD))
//   output (line l, (j/NS)*NS*RAD + (j mod NS) + r*NS)
// tw[m] = exp(-2 pi i m / N), m < N.  The arithmetic does not depend on the
// tile addressing XA.
template <int N, int RAD, int NS, int NL, int NTHR, bool INV, typename XA, typename V>
__device__ __forceinline__ void stockham_stage_at(const Tile<V> lds, const V *__restrict__ tw, int tid)
{
    constexpr int NBL = N / RAD;
    constexpr int NBF = NBL * NL;
    constexpr int PER = (NBF + NTHR - 1) / NTHR;
    V v[PER][RAD];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int bf = tid + q * NTHR;
        if ((NBF % NTHR == 0) || bf < NBF) {
            const int l = bf % NL, j = bf / NL;
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[q][r] = lds.ld(XA::template rd<NBL>(j, r, l));
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int bf = tid + q * NTHR;
        if ((NBF % NTHR == 0) || bf < NBF) {
            const int l = bf % NL, j = bf / NL;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r)
                    v[q][r] = cmulw<INV>(v[q][r], tw[r * k * (N / (NS * RAD))]);
            }
            dft<RAD, INV>(v[q]);
#pragma unroll
            for (int r = 0; r < RAD; ++r) lds.st(XA::template wr<NS, RAD>(j, r, l), v[q][r]);
        }
    }
    __syncthreads();
}

template <int N, int RAD, int NS, int NL, int NTHR, bool INV, typename V>
__device__ __forceinline__ void stockham_stage(const Tile<V> lds, const V *__restrict__ tw, int tid)
{
    stockham_stage_at<N, RAD, NS, NL, NTHR, INV, RowPad<NL>, V>(lds, tw, tid);
}

template <int N, int NS, int NL, int NTHR, bool INV, typename V>
__device__ __forceinline__ void fft_stages(const Tile<V> lds, const V *__restrict__ tw, int tid)
{
    if constexpr (NS < N) {
        constexpr int RAD = pick_radix(N / NS);
        static_assert(RAD == 2 || RAD == 3 || RAD == 4 || RAD == 5 || RAD == 8, "length must be 2^a 3^b 5^c");
        stockham_stage<N, RAD, NS, NL, NTHR, INV, V>(lds, tw, tid);
        fft_stages<N, NS * RAD, NL, NTHR, INV, V>(lds, tw, tid);
    }
}

// Stage counts of a length-m transform with pick_radix, and with a radix-16
// stage taken wherever it lowers that count (16, 48, 80, 128, 256; not 32,
// 64, 96, 512, whose stage count it leaves unchanged).
constexpr int stages_r8(int m) { return m <= 1 ? 0 : 1 + stages_r8(m / pick_radix(m)); }
constexpr int pick_radix_x(int m);
constexpr int stages_x(int m) { return m <= 1 ? 0 : 1 + stages_x(m / pick_radix_x(m)); }
constexpr int pick_radix_x(int m)
{
    return (m % 16 == 0 && 1 + stages_x(m / 16) < stages_r8(m)) ? 16 : pick_radix(m);
}

// The line passes offer both radix plans to the autotuner where they differ
// (lengths 16, 48, 80, 128, 256): plan 0 the radix-8 plan of fft_stages,
// plan 1 the radix-16 plan of fft_stages_x.
template <int N> constexpr bool radix16_plan_differs() { return stages_x(N) < stages_r8(N); }

// The same transform with those radix-16 stages (128 = 16 x 8: two LDS round
// trips instead of three).  Used by the x pass, whose tiles hold all S
// systems of a column: there the saved LDS traffic outweighs the 16 values a
// thread holds per butterfly (measured 220 -> 199 us FP64, 108 -> 102 us FP32
// at 128^3 x 12; at 512, where the stage count stays 3, radix 16 only cost
// registers).  The line passes keep radix 8.
template <int N, int NS, int NL, int NTHR, bool INV, typename XA = RowPad<NL>, typename V>
__device__ __forceinline__ void fft_stages_x(const Tile<V> lds, const V *__restrict__ tw, int tid)
{
    if constexpr (NS < N) {
        constexpr int RAD = pick_radix_x(N / NS);
        static_assert(RAD == 2 || RAD == 3 || RAD == 4 || RAD == 5 || RAD == 8 || RAD == 16,
                      "length must be 2^a 3^b 5^c");
        stockham_stage_at<N, RAD, NS, NL, NTHR, INV, XA, V>(lds, tw, tid);
        fft_stages_x<N, NS * RAD, NL, NTHR, INV, XA, V>(lds, tw, tid);
    }
}

// The stage of stockham_stage over a tile with arbitrary strides: element
// (line l, position p) at l * LS + p * PS.  Lanes stay line-fastest; the
// arithmetic (radix order, twiddle indices, butterflies) is that of
// stockham_stage, so a line transformed here is bit-identical to the same
// line transformed by fft_stages.  The plane kernels hold a whole y-z plane
// in one tile and transform it along both axes, once with unit line stride
// and once with unit position stride.
template <int N, int RAD, int NS, int NL, int LS, int PS, int NTHR, bool INV, typename V>
__device__ __forceinline__ void stockham_stage_g(const Tile<V> lds, const V *__restrict__ tw, int tid)
{
    constexpr int NBL = N / RAD;
    constexpr int NBF = NBL * NL;
    constexpr int PER = (NBF + NTHR - 1) / NTHR;
    V v[PER][RAD];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int bf = tid + q * NTHR;
        if ((NBF % NTHR == 0) || bf < NBF) {
            const int l = bf % NL, j = bf / NL;
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[q][r] = lds.ld(l * LS + (j + r * NBL) * PS);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int bf = tid + q * NTHR;
        if ((NBF % NTHR == 0) || bf < NBF) {
            const int l = bf % NL, j = bf / NL;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r)
                    v[q][r] = cmulw<INV>(v[q][r], tw[r * k * (N / (NS * RAD))]);
            }
            dft<RAD, INV>(v[q]);
            const int idxD = (j / NS) * NS * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) lds.st(l * LS + (idxD + r * NS) * PS, v[q][r]);
        }
    }
    __syncthreads();
}

template <int N, int NS, int NL, int LS, int PS, int NTHR, bool INV, typename V>
__device__ __forceinline__ void fft_stages_g(const Tile<V> lds, const V *__restrict__ tw, int tid)
{
    if constexpr (NS < N) {
        constexpr int RAD = pick_radix(N / NS);
        static_assert(RAD == 2 || RAD == 3 || RAD == 4 || RAD == 5 || RAD == 8, "length must be 2^a 3^b 5^c");
        stockham_stage_g<N, RAD, NS, NL, LS, PS, NTHR, INV, V>(lds, tw, tid);
        fft_stages_g<N, NS * RAD, NL, LS, PS, NTHR, INV, V>(lds, tw, tid);
    }
}

// Addressing of one line of a pass: line index L is split as (oa, ob, d),
// d fastest (d < D, ob < B); element p of the line sits at
//   oa*SA + ob*SB + d*SD + p*SP.
struct LineMap {
    long long SA, SB, SD, SP;
    long long D, B;
};

__device__ __forceinline__ long long line_base(const LineMap &m, long long L)
{
    const long long d = L % m.D, t = L / m.D;
    const long long ob = t % m.B, oa = t / m.B;
    return oa * m.SA + ob * m.SB + d * m.SD;
}

// Global -> LDS tile.  base[l] < 0 marks an inactive line.  Position-fastest
// when the line is contiguous in memory (SP == 1), line-fastest otherwise
// (then consecutive lines are contiguous: SD == 1).  Every load of the thread
// is issued before its first LDS write.  For 16-byte elements this beats a
// strided loop of 4 loads in flight on the 16-line tiles (FP64 forward passes
// at 128^3 x 12: 159.6 -> 136.0 us) and is within 2 % of it on the 8-line
// ones.
template <int N, int NL, int NTHR, typename V>
__device__ __forceinline__ void tile_load(const Tile<V> lds, const V *__restrict__ src,
                                          const long long *base, long long SP, int tid)
{
    constexpr int LDP = NL + 1, TOT = N * NL, PER = (TOT + NTHR - 1) / NTHR;
    {
        V r[PER];
        if (SP == 1) {
#pragma unroll
            for (int t = 0; t < PER; ++t) {
                const int e = tid + t * NTHR;
                const long long b = (TOT % NTHR == 0 || e < TOT) ? base[e / N] : -1;
                if (b >= 0) r[t] = ld_stream(&src[b + e % N]);
            }
#pragma unroll
            for (int t = 0; t < PER; ++t) {
                const int e = tid + t * NTHR;
                if (TOT % NTHR == 0 || e < TOT) lds.st((e % N) * LDP + e / N, r[t]);
            }
        } else {
#pragma unroll
            for (int t = 0; t < PER; ++t) {
                const int e = tid + t * NTHR;
                const long long b = (TOT % NTHR == 0 || e < TOT) ? base[e % NL] : -1;
                if (b >= 0) r[t] = ld_stream(&src[b + (e / NL) * SP]);
            }
#pragma unroll
            for (int t = 0; t < PER; ++t) {
                const int e = tid + t * NTHR;
                if (TOT % NTHR == 0 || e < TOT) lds.st((e / NL) * LDP + e % NL, r[t]);
            }
        }
    }
}

template <int N, int NL, int NTHR, typename V>
__device__ __forceinline__ void tile_store(V *__restrict__ dst, const Tile<V> lds,
                                           const long long *base, long long SP, int tid)
{
    constexpr int LDP = NL + 1, TOT = N * NL;
    if (SP == 1) {
#pragma unroll 4
        for (int e = tid; e < TOT; e += NTHR) {
            const int l = e / N, p = e % N;
            const long long b = base[l];
            if (b >= 0) st_stream(&dst[b + p], lds.ld(p * LDP + l));
        }
    } else {
#pragma unroll 4
        for (int e = tid; e < TOT; e += NTHR) {
            const int l = e % NL, p = e / NL;
            const long long b = base[l];
            if (b >= 0) st_stream(&dst[b + p * SP], lds.ld(p * LDP + l));
        }
    }
}

// ------------------------------------------------------------- reductions
// Compensated (Kahan) running sum of doubles, for sums that run over many
// terms in one thread (k-points of the x pass, planes of the plane update).
struct KSum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x)
    {
        const double y = x - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// Error-free sum a + b = s + e (Knuth's TwoSum, no ordering condition).
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Block-wide sum of NV doubles (NTHR threads, wave64).  Result valid in
// thread 0.  Uses `red` (NV * NTHR/64 doubles of LDS).
template <int NV, int NTHR>
__device__ __forceinline__ void block_sum(double *v, double *red, int tid)
{
    constexpr int NW = NTHR / 64;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[q] = x;
    }
    const int lane = tid & 63, w = tid >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q * NW + w] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int i = 0; i < NW; ++i) s += red[q * NW + i];
            v[q] = s;
        }
    }
}

// Block-wide sum of NV doubles (NTHR threads, wave64) for the energy and
// average reductions: a pairwise shuffle tree within each wave (6 levels,
// relative error at most 6 u for terms of one sign), then the wave sums
// added with TwoSum and their rounding errors added back.  Compensating every
// shuffle level as well costs as much as 8 % of the update kernel (measured
// at 128^3 x 12) for no visible gain.  Result valid in thread 0.  Uses `red`
// (NV * NTHR/64 doubles of LDS).
template <int NV, int NTHR>
__device__ __forceinline__ void block_sum_comp(double *v, double *red, int tid)
{
    constexpr int NW = NTHR / 64;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double x = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[q] = x;
    }
    const int lane = tid & 63, w = tid >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) red[q * NW + w] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0, e = 0.0;
            for (int i = 0; i < NW; ++i) {
                double t, r;
                two_sum(s, red[q * NW + i], t, r);
                s = t;
                e += r;
            }
            v[q] = s + e;
        }
    }
}

} // namespace pfdd
