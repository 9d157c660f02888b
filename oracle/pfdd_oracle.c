/*
 * pfdd_oracle.c -- CPU FP64 restatement of PFDD-2.x's Ginzburg-Landau (GL) step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker.  The product library never links or calls it.
 *
 * Every function restates one piece of the reference (file:line into the
 * lanl/Phase-Field-Dislocation-Dynamics-PFDD-2.x source tree, src/ directory).
 * The reference itself cannot be built in this image (it needs FFTW 2.1.5 MPI
 * and csh-generated style headers), so each function is pinned by independent
 * checks in tests/test_oracle.py (numpy FFT, naive DFT, closed-form Green's
 * function, finite differences of the energy, invariants).
 *
 * Data layout is the reference's:
 *   xi[2*(k + j*nz + i*ny*nz + a*N) + {0,1}]   (re, im), N = nx*ny*nz
 *   BB[r + a*N + b*N*S]                         (fft_fftw_slab.cpp:631)
 *   FF/DD[r + a*N + i*N*S + j*N*S*3]            (fft_fftw_slab.cpp:725-729)
 *
 * Parity status of each function: see the header comment of each.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DELTA(i, j) ((i) == (j) ? 1.0 : 0.0)

static int g_threads = 1;

void orc_set_threads(int n)
{
#ifdef _OPENMP
    g_threads = n > 0 ? n : 1;
#else
    (void)n;
    g_threads = 1;
#endif
}

int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* Material constants: material.cpp:156-167                                  */
/* out: C44, C12, C11, nu, S11, S12, S44, ll                                 */
/* ------------------------------------------------------------------------ */
void orc_material(double mu, double young, double *out)
{
    double C44 = mu;
    double nu = young / 2.0 / (mu)-1.0;
    double C12 = 2.0 * nu * C44 / (1.0 - 2.0 * nu);
    double C11 = 2.0 * C44 + C12;
    double S11 = 1.0 / (young);
    double S12 = -nu / (young);
    double S44 = 2 * (S11 - S12);
    out[0] = C44; out[1] = C12; out[2] = C11; out[3] = nu;
    out[4] = S11; out[5] = S12; out[6] = S44; out[7] = C12;
}

/* Voigt re-derivation of the isotropic moduli done inside Bmatrix
 * (fft_fftw_slab.cpp:487-490) and stress() (:3497-3500).
 * out: mu, ll, young, xnu */
void orc_voigt(double C44, double C12, double C11, double *out)
{
    double mu = C44 - (2.0 * C44 + C12 - C11) / 5.0;
    double ll = C12 - (2.0 * C44 + C12 - C11) / 5.0;
    double young = mu * (3 * ll + 2 * mu) / (ll + mu);
    double xnu = young / 2.0 / mu - 1.0;
    out[0] = mu; out[1] = ll; out[2] = young; out[3] = xnu;
}

/* ------------------------------------------------------------------------ */
/* Primitive vectors: column normalisation (fft.cpp:101-133) and inverse    */
/* (fft.cpp:155-169).  prim[3*r+c] row-major, columns are the x,y,z vectors. */
/* ------------------------------------------------------------------------ */
void orc_prim(const double *x, const double *y, const double *z, double *prim,
              double *invprim, double *norvol)
{
    const double *v[3] = {x, y, z};
    double p[3][3], ip[3][3];
    for (int c = 0; c < 3; c++) {
        double n = sqrt(v[c][0] * v[c][0] + v[c][1] * v[c][1] + v[c][2] * v[c][2]);
        for (int r = 0; r < 3; r++) p[r][c] = v[c][r] / n;
    }
    double nv = p[1][0] * p[2][1] * p[0][2] - p[2][0] * p[1][1] * p[0][2] +
                p[2][0] * p[0][1] * p[1][2] - p[0][0] * p[2][1] * p[1][2] +
                p[0][0] * p[1][1] * p[2][2] - p[1][0] * p[0][1] * p[2][2];
    ip[0][0] = (p[1][1] * p[2][2] - p[2][1] * p[1][2]) / nv;
    ip[0][1] = (p[2][1] * p[0][2] - p[0][1] * p[2][2]) / nv;
    ip[0][2] = (p[0][1] * p[1][2] - p[1][1] * p[0][2]) / nv;
    ip[1][0] = (p[1][2] * p[2][0] - p[2][2] * p[1][0]) / nv;
    ip[1][1] = (p[2][2] * p[0][0] - p[0][2] * p[2][0]) / nv;
    ip[1][2] = (p[0][2] * p[1][0] - p[1][2] * p[0][0]) / nv;
    ip[2][0] = (p[1][0] * p[2][1] - p[2][0] * p[1][1]) / nv;
    ip[2][1] = (p[2][0] * p[0][1] - p[0][0] * p[2][1]) / nv;
    ip[2][2] = (p[0][0] * p[1][1] - p[1][0] * p[0][1]) / nv;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            prim[3 * r + c] = p[r][c];
            invprim[3 * r + c] = ip[r][c];
        }
    *norvol = nv;
}

/* ------------------------------------------------------------------------ */
/* Reciprocal matrix: App::set_scale, app.cpp:383-431.                       */
/* ------------------------------------------------------------------------ */
void orc_set_scale(const double *prim, double sa, double sb, double sc,
                   double *four, double *normfour, double *vol_out)
{
    double s[3][3], f[3][3];
    for (int r = 0; r < 3; r++) {
        s[r][0] = sa * prim[3 * r + 0];
        s[r][1] = sb * prim[3 * r + 1];
        s[r][2] = sc * prim[3 * r + 2];
    }
    double vol = s[1][0] * s[2][1] * s[0][2] - s[2][0] * s[1][1] * s[0][2] +
                 s[2][0] * s[0][1] * s[1][2] - s[0][0] * s[2][1] * s[1][2] +
                 s[0][0] * s[1][1] * s[2][2] - s[1][0] * s[0][1] * s[2][2];
    f[0][0] = (s[1][1] * s[2][2] - s[2][1] * s[1][2]) / vol;
    f[0][1] = (s[2][1] * s[0][2] - s[0][1] * s[2][2]) / vol;
    f[0][2] = (s[0][1] * s[1][2] - s[1][1] * s[0][2]) / vol;
    f[1][0] = (s[1][2] * s[2][0] - s[2][2] * s[1][0]) / vol;
    f[1][1] = (s[2][2] * s[0][0] - s[0][2] * s[2][0]) / vol;
    f[1][2] = (s[0][2] * s[1][0] - s[1][2] * s[0][0]) / vol;
    f[2][0] = (s[1][0] * s[2][1] - s[2][0] * s[1][1]) / vol;
    f[2][1] = (s[2][0] * s[0][1] - s[0][0] * s[2][1]) / vol;
    f[2][2] = (s[0][0] * s[1][1] - s[1][0] * s[0][1]) / vol;
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) four[3 * r + c] = f[r][c];
        normfour[r] = sqrt(f[r][0] * f[r][0] + f[r][1] * f[r][1] + f[r][2] * f[r][2]);
    }
    *vol_out = vol;
}

/* Slip-vector rotation into the grid frame: AppFCC::set_slip rotation block,
 * app_fcc.cpp:288-313 (identical in app_3d_hcp_basal.cpp:248-267). */
void orc_rotate_slip(int ns, double *xn, double *xb, const double *four,
                     const double *normfour)
{
    for (int i = 0; i < ns; i++) {
        double *v[2] = {xn + 3 * i, xb + 3 * i};
        for (int w = 0; w < 2; w++) {
            double a = (v[w][0] * four[0] + v[w][1] * four[1] + v[w][2] * four[2]) / normfour[0];
            double b = (v[w][0] * four[3] + v[w][1] * four[4] + v[w][2] * four[5]) / normfour[1];
            double c = (v[w][0] * four[6] + v[w][1] * four[7] + v[w][2] * four[8]) / normfour[2];
            v[w][0] = a; v[w][1] = b; v[w][2] = c;
        }
    }
}

/* sigma_rot = prim sigma prim^-1 if |det prim| in (0.99,1.01), else sigma:
 * FFTW_Slab::rotate_stress, fft_fftw_slab.cpp:242-274. */
void orc_rotate_stress(const double *prim, const double *invprim, double norvol,
                       const double *sigma, double *sigma_rot)
{
    int ortho = (fabs(norvol) < 1.01) && (fabs(norvol) > 0.99);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            if (ortho) {
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++)
                        s += prim[3 * i + k] * sigma[3 * k + l] * invprim[3 * l + j];
            } else {
                s = sigma[3 * i + j];
            }
            sigma_rot[3 * i + j] = s;
        }
}

/* tau_a = sum_ij sigma_rot_ij n_aj b_ai / dslip: resolSS_Schmid, :737-769. */
void orc_resolss(int ns, const double *sigma_rot, const double *xn, const double *xb,
                 double dslip, double *tau)
{
    for (int is = 0; is < ns; is++) {
        double t = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                t = t + (sigma_rot[3 * i + j] * xn[3 * is + j] * xb[3 * is + i]) / dslip;
        tau[is] = t;
    }
}

/* ------------------------------------------------------------------------ */
/* Frequencies: FFTW_Slab::frec, fft_fftw_slab.cpp:280-335 (one rank:        */
/* local_x = nx, local_x_start = 0).                                         */
/* ------------------------------------------------------------------------ */
static double frac_freq(int kk, int n)
{
    if (kk >= n / 2) return (double)(kk - n) / (double)n;
    return (double)kk / (double)n;
}

void orc_frec(int nx, int ny, int nz, const double *four, double *fx, double *fy, double *fz)
{
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long nf = k + (long)j * nz + (long)i * nz * ny;
                double kx = frac_freq(i, nx), ky = frac_freq(j, ny), kz = frac_freq(k, nz);
                fx[nf] = (kx * four[0] + ky * four[3] + kz * four[6]);
                fy[nf] = (kx * four[1] + ky * four[4] + kz * four[7]);
                fz[nf] = (kx * four[2] + ky * four[5] + kz * four[8]);
            }
}

/* Isotropic stiffness C_ijkl = C44 (d_ik d_jm + d_im d_jk) + C12 d_ij d_km:
 * FFTW_Slab::stiffness, fft_fftw_slab.cpp:373-392.  C[27*i+9*j+3*k+m]. */
void orc_stiffness(double C44, double C12, double *C)
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++)
                for (int m = 0; m < 3; m++)
                    C[27 * i + 9 * j + 3 * k + m] =
                        C44 * (DELTA(i, k) * DELTA(j, m) + DELTA(i, m) * DELTA(j, k)) +
                        C12 * DELTA(i, j) * DELTA(k, m);
}

/* ------------------------------------------------------------------------ */
/* B(k): FFTW_Slab::Bmatrix, fft_fftw_slab.cpp:466-638, restated literally   */
/* (A-tensor route).  BB[r + a*N + b*N*S] = B_ab(k_r)/mu.  eps_out[9*a+3i+j] */
/* receives eps_a = b_a n_a^T / dslip (:546-553).  moduli_out = mu,ll,E,nu.  */
/* As in the reference, A keeps its zero initialisation at f = 0 because     */
/* k = 0 is the first point visited, so B(0) = 0.                            */
/* ------------------------------------------------------------------------ */
void orc_bmatrix(int nx, int ny, int nz, int ns, double C44, double C12, double C11,
                 double dslip, const double *xn, const double *xb, const double *fx,
                 const double *fy, const double *fz, double *BB, double *eps_out,
                 double *moduli_out)
{
    long N = (long)nx * ny * nz;
    double C[81];
    orc_stiffness(C44, C12, C);
    double mod[4];
    orc_voigt(C44, C12, C11, mod);
    double mu = mod[0], xnu = mod[3];
    if (moduli_out) memcpy(moduli_out, mod, sizeof(mod));
    double *eps = (double *)malloc(sizeof(double) * 9 * ns);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < ns; k++)
                eps[9 * k + 3 * i + j] = xb[3 * k + i] * xn[3 * k + j] / dslip;
    if (eps_out) memcpy(eps_out, eps, sizeof(double) * 9 * ns);

#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long nfreq = 0; nfreq < N; nfreq++) {
        double A[81];
        double G[9];
        double fk[3] = {fx[nfreq], fy[nfreq], fz[nfreq]};
        double fk2 = fk[0] * fk[0] + fk[1] * fk[1] + fk[2] * fk[2];
        double fk4 = fk2 * fk2;
        for (int q = 0; q < 81; q++) A[q] = 0.0;
        if (fk2 > 0) {
            for (int m = 0; m < 3; m++)
                for (int n = 0; n < 3; n++)
                    for (int u = 0; u < 3; u++)
                        for (int v = 0; v < 3; v++) {
                            double a = 0.0;
                            for (int i = 0; i < 3; i++)
                                for (int j = 0; j < 3; j++)
                                    for (int k = 0; k < 3; k++) {
                                        G[3 * k + i] = (2.0 * DELTA(i, k) / fk2 -
                                                        1.0 / (1.0 - xnu) * fk[i] * fk[k] / fk4) /
                                                       (2.0 * C44);
                                        for (int l = 0; l < 3; l++)
                                            a = a - C[27 * k + 9 * l + 3 * u + v] *
                                                        C[27 * i + 9 * j + 3 * m + n] *
                                                        G[3 * k + i] * fk[j] * fk[l];
                                    }
                            A[27 * m + 9 * n + 3 * u + v] = a + C[27 * m + 9 * n + 3 * u + v];
                        }
        }
        for (int ka = 0; ka < ns; ka++)
            for (int kb = 0; kb < ns; kb++) {
                double B = 0.0;
                for (int m = 0; m < 3; m++)
                    for (int n = 0; n < 3; n++)
                        for (int u = 0; u < 3; u++)
                            for (int v = 0; v < 3; v++)
                                B = B + A[27 * m + 9 * n + 3 * u + v] * eps[9 * ka + 3 * m + n] *
                                            eps[9 * kb + 3 * u + v];
                BB[nfreq + ka * N + kb * N * ns] = B / mu;
            }
    }
    free(eps);
}

/* ------------------------------------------------------------------------ */
/* F(k), D(k): FFTW_Slab::Fmatrix, fft_fftw_slab.cpp:644-732.                */
/* mu, xnu are the Voigt moduli stored by Bmatrix (material->mu, ->xnu).     */
/* The reference's G persists across k; at k = 0 (first point) it is zero.   */
/* ------------------------------------------------------------------------ */
void orc_fmatrix(int nx, int ny, int nz, int ns, double C44, double C12, double mu,
                 double xnu, const double *eps, const double *fx, const double *fy,
                 const double *fz, double *FF, double *DD)
{
    long N = (long)nx * ny * nz;
    double C[81];
    orc_stiffness(C44, C12, C);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long nfreq = 0; nfreq < N; nfreq++) {
        double A[81], G[9];
        for (int q = 0; q < 81; q++) A[q] = 0.0;
        for (int q = 0; q < 9; q++) G[q] = 0.0;
        double fk[3] = {fx[nfreq], fy[nfreq], fz[nfreq]};
        double fk2 = fk[0] * fk[0] + fk[1] * fk[1] + fk[2] * fk[2];
        double fk4 = fk2 * fk2;
        if (fk2 > 0) {
            for (int m = 0; m < 3; m++)
                for (int n = 0; n < 3; n++)
                    for (int u = 0; u < 3; u++)
                        for (int v = 0; v < 3; v++) {
                            double a = 0.0;
                            for (int i = 0; i < 3; i++)
                                for (int j = 0; j < 3; j++)
                                    for (int k = 0; k < 3; k++) {
                                        G[3 * k + i] = (2.0 * DELTA(i, k) / fk2 -
                                                        1.0 / (1.0 - xnu) * fk[i] * fk[k] / fk4) /
                                                       (2.0 * mu);
                                        for (int l = 0; l < 3; l++)
                                            a = a - C[27 * k + 9 * l + 3 * u + v] *
                                                        C[27 * i + 9 * j + 3 * m + n] *
                                                        G[3 * k + i] * fk[j] * fk[l];
                                    }
                            A[27 * m + 9 * n + 3 * u + v] = a + C[27 * m + 9 * n + 3 * u + v];
                        }
        }
        for (int ka = 0; ka < ns; ka++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double F = 0.0, D = 0.0;
                    for (int k = 0; k < 3; k++)
                        for (int l = 0; l < 3; l++) {
                            for (int m = 0; m < 3; m++)
                                for (int n = 0; n < 3; n++)
                                    F = F + C[27 * k + 9 * l + 3 * m + n] * G[3 * j + k] *
                                                eps[9 * ka + 3 * m + n] * fk[i] * fk[l];
                            D = D + A[27 * i + 9 * j + 3 * k + l] * eps[9 * ka + 3 * k + l];
                        }
                    long nb = nfreq + ka * N + (long)i * N * ns + (long)j * N * ns * 3;
                    FF[nb] = F;
                    if (DD) DD[nb] = D;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* FFT with FFTW 2 semantics (in place, unnormalised, forward sign -1,       */
/* backward sign +1: the fftwnd_mpi calls at fft_fftw_slab.cpp:1962,2021).  */
/* Mixed-radix recursive decimation in time; any n.                          */
/* ------------------------------------------------------------------------ */
static void dft_rec(int n, int n0, long stride, const double *in, double *out,
                    const double *tw /* e^{sign 2 pi i j / n0}, 2*n0 */)
{
    if (n == 1) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    int p;
    if (n % 4 == 0) p = 4;
    else if (n % 2 == 0) p = 2;
    else if (n % 3 == 0) p = 3;
    else if (n % 5 == 0) p = 5;
    else {
        p = n;
        for (int d = 7; d * d <= n; d += 2)
            if (n % d == 0) { p = d; break; }
    }
    int m = n / p;
    for (int q = 0; q < p; q++) dft_rec(m, n0, stride * p, in + 2 * q * stride, out + 2 * q * m, tw);
    int step = n0 / n;          /* W_n = tw[step] */
    int pstep = n0 / p;         /* W_p = tw[pstep] */
    double t[2 * 64];
    double *tb = p <= 64 ? t : (double *)malloc(sizeof(double) * 2 * p);
    double *ob = (double *)malloc(sizeof(double) * 2 * p);
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            long e = ((long)q * k * step) % n0;
            double wr = tw[2 * e], wi = tw[2 * e + 1];
            double yr = out[2 * (q * m + k)], yi = out[2 * (q * m + k) + 1];
            tb[2 * q] = yr * wr - yi * wi;
            tb[2 * q + 1] = yr * wi + yi * wr;
        }
        for (int s = 0; s < p; s++) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < p; q++) {
                long e = ((long)q * s * pstep) % n0;
                double wr = tw[2 * e], wi = tw[2 * e + 1];
                sr += tb[2 * q] * wr - tb[2 * q + 1] * wi;
                si += tb[2 * q] * wi + tb[2 * q + 1] * wr;
            }
            ob[2 * s] = sr;
            ob[2 * s + 1] = si;
        }
        for (int s = 0; s < p; s++) {
            out[2 * (k + s * m)] = ob[2 * s];
            out[2 * (k + s * m) + 1] = ob[2 * s + 1];
        }
    }
    free(ob);
    if (tb != t) free(tb);
}

static double *make_twiddles(int n, int sign)
{
    double *tw = (double *)malloc(sizeof(double) * 2 * n);
    for (int j = 0; j < n; j++) {
        /* reduce the angle to [0, pi/4]-style symmetric evaluation for accuracy */
        double ang = 2.0 * M_PI * (double)j / (double)n;
        tw[2 * j] = cos(ang);
        tw[2 * j + 1] = sign * sin(ang);
    }
    /* exact values at the quarter points */
    if (n % 4 == 0) {
        tw[2 * (n / 4)] = 0.0; tw[2 * (n / 4) + 1] = sign * 1.0;
        tw[2 * (3 * n / 4)] = 0.0; tw[2 * (3 * n / 4) + 1] = -sign * 1.0;
    }
    if (n % 2 == 0) { tw[2 * (n / 2)] = -1.0; tw[2 * (n / 2) + 1] = 0.0; }
    return tw;
}

/* 1-D transform of `howmany` lines of length n, element stride `stride`,
 * line starts given by base + line_offset(l). */
static void fft_axis(double *data, int n0, int n1, int n2, int axis, int sign)
{
    int dims[3] = {n0, n1, n2};
    int n = dims[axis];
    if (n == 1) return;
    long strides[3] = {(long)n1 * n2, n2, 1};
    long stride = strides[axis];
    int oa = (axis + 1) % 3, ob = (axis + 2) % 3;
    long nlines = (long)dims[oa] * dims[ob];
    double *tw = make_twiddles(n, sign);
#pragma omp parallel num_threads(g_threads)
    {
        double *bi = (double *)malloc(sizeof(double) * 2 * n);
        double *bo = (double *)malloc(sizeof(double) * 2 * n);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; l++) {
            long ia = l / dims[ob], ib = l % dims[ob];
            long base = ia * strides[oa] + ib * strides[ob];
            for (int t = 0; t < n; t++) {
                bi[2 * t] = data[2 * (base + t * stride)];
                bi[2 * t + 1] = data[2 * (base + t * stride) + 1];
            }
            dft_rec(n, n, 1, bi, bo, tw);
            for (int t = 0; t < n; t++) {
                data[2 * (base + t * stride)] = bo[2 * t];
                data[2 * (base + t * stride) + 1] = bo[2 * t + 1];
            }
        }
        free(bi);
        free(bo);
    }
    free(tw);
}

/* In-place unnormalised 3-D DFT of an n0 x n1 x n2 complex array (last index
 * fastest).  sign = -1 forward (FFTW_FORWARD), +1 backward. */
void orc_fft3d(int n0, int n1, int n2, double *data, int sign)
{
    fft_axis(data, n0, n1, n2, 2, sign);
    fft_axis(data, n0, n1, n2, 1, sign);
    fft_axis(data, n0, n1, n2, 0, sign);
}

/* Plain O(n^2) DFT along one line: the definition, used to pin orc_fft3d. */
void orc_dft1d_naive(int n, const double *in, double *out, int sign)
{
    for (int k = 0; k < n; k++) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            long e = ((long)j * k) % n;
            long double ang = 2.0L * 3.14159265358979323846264338327950288L * e / n;
            long double c = cosl(ang), s = sign * sinl(ang);
            sr += in[2 * j] * c - in[2 * j + 1] * s;
            si += in[2 * j] * s + in[2 * j + 1] * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ------------------------------------------------------------------------ */
/* Core energy.                                                              */
/*  flag 1 core_energy_perfect :1130-1191 : par[0] = An' = An/(mu dslip b)    */
/*  flag 3 core_energy_sine    :1360-1418 : par[0] = isf', par[1] = Cn'      */
/*  flag 7 core_energy_mpea    :1196-1235 : amp[r + a N] = A/(mu dslip b)    */
/*  flag 8 (sine series; generalises 1/3/7): par[a*nm + m-1] = c_{a,m},       */
/*         optional amp; density amp * sum_m c_m sin^2(m pi x)               */
/* dE[r + a N] is the derivative with respect to Re xi (the force term of    */
/* update_order_parameter).  Returns E_core as the reference accumulates it  */
/* (sum / nx on one rank, quirk of :1407-1414); *e_var receives the plain    */
/* sum of the energy density (the variational core energy).                  */
/* ------------------------------------------------------------------------ */
double orc_core_energy(int nx, int ny, int nz, int ns, const double *xi, int flag,
                       const double *par, int nm, const double *amp, double *dE,
                       double *e_var)
{
    long N = (long)nx * ny * nz;
    double E_ref = 0.0, E_var = 0.0;
    for (int isa = 0; isa < ns; isa++)
        for (long r = 0; r < N; r++) {
            long index = r + isa * N;
            double x = xi[2 * index];
            double d = 0.0, ed = 0.0;
            if (flag == 1) {
                double An = par[0];
                ed = An * (sin(M_PI * x) * sin(M_PI * x));
                E_ref += ed / nx;
                d = An * M_PI * sin(2.0 * M_PI * x);
            } else if (flag == 3) {
                double isf = par[0], Cn = par[1];
                ed = (isf * (sin(M_PI * x) * sin(M_PI * x)) + Cn * (sin(2 * M_PI * x) * sin(2 * M_PI * x)));
                E_ref += ed / nx;
                d = (isf * M_PI * sin(2 * M_PI * x) + Cn * 2 * M_PI * sin(4 * M_PI * x));
            } else if (flag == 7) {
                double A = amp[index];
                ed = A * (sin(M_PI * x) * sin(M_PI * x));
                E_ref += ed / nx;
                d = A * M_PI * sin(2.0 * M_PI * x);
            } else if (flag == 5) {
                /* core_energy_bcc_perfect, :1071-1125: usf' scaled by a
                 * quadratic in the character angle theta (passed in amp[]) */
                double th = amp[index], cof;
                if (th < 0.0) cof = 1.0;
                else if (th < 1.23) cof = 1.1603 * th * th - 2.0431 * th + 1;
                else cof = 0.5473 * th * th - 2.0035 * th + 1.8923;
                ed = par[0] * cof * (sin(M_PI * x) * sin(M_PI * x));
                E_ref += ed / nx;
                d = par[0] * cof * M_PI * sin(2.0 * M_PI * x);
            } else if (flag == 4 || flag == 9) {
                /* core_energy_pyrII, :1429-1493: Fourier series
                 * par = aa0..aa4, bb1..bb4 (already / (mu dslip b)).  Flag 4
                 * is the routine as executed: its dE_core assignment sits after
                 * the cell loop (:1491), so only the last cell of each system
                 * gets the derivative (init_loop zeroes the rest) and E_core is
                 * never accumulated.  Flag 9 applies the derivative to every
                 * cell (the Fourier-series model the routine describes) and
                 * accumulates E_core as flags 1/3/7 do. */
                double c[5] = {0, 0, 0, 0, 0}, sn[5] = {0, 0, 0, 0, 0};
                for (int m = 1; m <= 4; m++) {
                    c[m] = cos(2.0 * m * M_PI * x);
                    sn[m] = sin(2.0 * m * M_PI * x);
                }
                ed = par[0];
                for (int m = 1; m <= 4; m++) ed += par[m] * c[m] + par[4 + m] * sn[m];
                double dd = 0.0;
                for (int m = 1; m <= 4; m++) dd += (2.0 * m * M_PI) * (par[4 + m] * c[m] - par[m] * sn[m]);
                d = (flag == 9 || r == N - 1) ? dd : 0.0;
                if (flag == 9) E_ref += ed / nx;
            } else if (flag == 8) {
                double s = 0.0, e = 0.0;
                for (int m = 1; m <= nm; m++) {
                    double c = par[isa * nm + m - 1];
                    double sm = sin(m * M_PI * x);
                    e += c * sm * sm;
                    s += c * (m * M_PI) * sin(2.0 * m * M_PI * x);
                }
                double A = amp ? amp[index] : 1.0;
                ed = A * e;
                E_ref += ed / nx;
                d = A * s;
            }
            dE[index] = d;
            E_var += ed;
        }
    if (flag == 2) {
        /* core_energy_extended, :1243-1354: coupled gamma surface of the three
         * systems of each plane (systems 3p, 3p+1, 3p+2); par = c0..c4, a1, a3
         * (already / (mu dslip b)).  dE of system j is the derivative of the
         * plane's density with respect to its Re xi, transcribed term by term
         * from df1core/df2core/df3core; E_core is not accumulated (:1320). */
        const double c0 = par[0], c1 = par[1], c2 = par[2], c3 = par[3], c4 = par[4], a1 = par[5], a3 = par[6];
        const double tp = 2.0 * M_PI;
        E_var = 0.0;
        for (int pl = 0; pl < ns / 3; pl++)
            for (long r = 0; r < N; r++) {
                const long i1 = r + (3 * pl) * N, i2 = i1 + N, i3 = i2 + N;
                const double x1 = xi[2 * i1], x2 = xi[2 * i2], x3 = xi[2 * i3];
                const double f =
                    c0 + c1 * (cos(tp * (x1 - x2)) + cos(tp * (x2 - x3)) + cos(tp * (x3 - x1))) +
                    c2 * (cos(tp * (2 * x1 - x2 - x3)) + cos(tp * (2 * x2 - x3 - x1)) + cos(tp * (2 * x3 - x1 - x2))) +
                    c3 * (cos(2 * tp * (x1 - x2)) + cos(2 * tp * (x2 - x3)) + cos(2 * tp * (x3 - x1))) +
                    c4 * (cos(tp * (3 * x1 - x2 - 2 * x3)) + cos(tp * (3 * x1 - 2 * x2 - x3)) +
                          cos(tp * (3 * x2 - x3 - 2 * x1)) + cos(tp * (3 * x2 - 2 * x3 - x1)) +
                          cos(tp * (3 * x3 - x1 - 2 * x2)) + cos(tp * (3 * x3 - 2 * x1 - x2))) +
                    a1 * (sin(tp * (x1 - x2)) + sin(tp * (x2 - x3)) + sin(tp * (x3 - x1))) +
                    a3 * (sin(2 * tp * (x1 - x2)) + sin(2 * tp * (x2 - x3)) + sin(2 * tp * (x3 - x1)));
                E_var += f;
                dE[i1] = tp * (c1 * (sin(tp * (x2 - x1)) + sin(tp * (x3 - x1))) +
                               c2 * (2.0 * sin(tp * (x2 + x3 - 2 * x1)) + sin(tp * (2 * x2 - x3 - x1)) +
                                     sin(tp * (2 * x3 - x1 - x2))) +
                               2.0 * c3 * (sin(2 * tp * (x2 - x1)) + sin(2 * tp * (x3 - x1))) +
                               c4 * (3.0 * sin(tp * (x2 + 2 * x3 - 3 * x1)) + 3.0 * sin(tp * (2 * x2 + x3 - 3 * x1)) +
                                     2.0 * sin(tp * (3 * x2 - x3 - 2 * x1)) + sin(tp * (3 * x2 - 2 * x3 - x1)) +
                                     sin(tp * (3 * x3 - x1 - 2 * x2)) + 2.0 * sin(tp * (3 * x3 - 2 * x1 - x2))) +
                               a1 * (cos(tp * (x1 - x2)) - cos(tp * (x3 - x1))) +
                               2.0 * a3 * (cos(2 * tp * (x1 - x2)) - cos(2 * tp * (x3 - x1))));
                dE[i2] = tp * (c1 * (sin(tp * (x1 - x2)) + sin(tp * (x3 - x2))) +
                               c2 * (sin(tp * (2 * x1 - x2 - x3)) + 2.0 * sin(tp * (x3 + x1 - 2 * x2)) +
                                     sin(tp * (2 * x3 - x1 - x2))) +
                               2.0 * c3 * (sin(2 * tp * (x1 - x2)) + sin(2 * tp * (x3 - x2))) +
                               c4 * (sin(tp * (3 * x1 - x2 - 2 * x3)) + 2.0 * sin(tp * (3 * x1 - 2 * x2 - x3)) +
                                     3.0 * sin(tp * (x3 + 2 * x1 - 3 * x2)) + 3.0 * sin(tp * (2 * x3 + x1 - 3 * x2)) +
                                     2.0 * sin(tp * (3 * x3 - x1 - 2 * x2)) + sin(tp * (3 * x3 - 2 * x1 - x2))) +
                               a1 * (cos(tp * (x2 - x3)) - cos(tp * (x1 - x2))) +
                               2.0 * a3 * (cos(2 * tp * (x2 - x3)) - cos(2 * tp * (x1 - x2))));
                dE[i3] = tp * (c1 * (sin(tp * (x2 - x3)) + sin(tp * (x1 - x3))) +
                               c2 * (sin(tp * (2 * x1 - x2 - x3)) + sin(tp * (2 * x2 - x3 - x1)) +
                                     2.0 * sin(tp * (x1 + x2 - 2 * x3))) +
                               2.0 * c3 * (sin(2 * tp * (x2 - x3)) + sin(2 * tp * (x1 - x3))) +
                               c4 * (2.0 * sin(tp * (3 * x1 - x2 - 2 * x3)) + sin(tp * (3 * x1 - 2 * x2 - x3)) +
                                     sin(tp * (3 * x2 - x3 - 2 * x1)) + 2.0 * sin(tp * (3 * x2 - 2 * x3 - x1)) +
                                     3.0 * sin(tp * (x1 + 2 * x2 - 3 * x3)) + 3.0 * sin(tp * (2 * x1 + x2 - 3 * x3))) +
                               a1 * (cos(tp * (x3 - x1)) - cos(tp * (x2 - x3))) +
                               2.0 * a3 * (cos(2 * tp * (x3 - x1)) - cos(2 * tp * (x2 - x3))));
            }
        E_ref = 0.0;
    } else if (flag == 4) {
        E_ref = 0.0;
    }
    if (e_var) *e_var = E_var;
    return E_ref;
}

/* resolSS_non_Schmid, :774-845, for ns systems with Burgers vectors xb
 * (rows of 3): MRSS plane n_is at (angle + 120 is) deg about y, n'_is at
 * +60 deg; tau = sum_ij sigma_rot_ij S_ij / dslip with
 * S_ij = n_j b_i + w1 n'_j b_i + w2 n_j (n x b)_i + w3 n'_j (n' x b)_i. */
void orc_resolss_non_schmid(int ns, const double *xb, const double *sigma_rot, double dslip, double angle,
                            double w1, double w2, double w3, double *tau)
{
    double theta_n = 0.0;
    for (int is = 0; is < ns; is++) {
        double xm[3], xp[3], nxb[3], npxb[3];
        const double *b = xb + 3 * is;
        xm[0] = -sin((theta_n + angle) * M_PI / 180.);
        xm[1] = 0.0;
        xm[2] = cos((theta_n + angle) * M_PI / 180.);
        xp[0] = -sin((theta_n + angle + 60) * M_PI / 180.);
        xp[1] = 0.0;
        xp[2] = cos((theta_n + angle + 60) * M_PI / 180.);
        theta_n += 120.;
        nxb[0] = xm[1] * b[2] - xm[2] * b[1];
        nxb[1] = -xm[0] * b[2] + xm[2] * b[0];
        nxb[2] = xm[0] * b[1] - xm[1] * b[0];
        npxb[0] = xp[1] * b[2] - xp[2] * b[1];
        npxb[1] = -xp[0] * b[2] + xp[2] * b[0];
        npxb[2] = xp[0] * b[1] - xp[1] * b[0];
        tau[is] = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double S = xm[j] * b[i] + w1 * xp[j] * b[i] + w2 * xm[j] * nxb[i] + w3 * xp[j] * npxb[i];
                tau[is] = tau[is] + (sigma_rot[3 * i + j] * S) / dslip;
            }
    }
}

/* gradient(), :869-1065, one rank: periodic central differences of Re xi,
 * transformed to the global axes with the reference's in-place order (the
 * y and z components use the already transformed x (and y) component, quirk
 * Q6), tang = grad x n, theta = acos(tang . b) flipped to pi - theta when
 * (b x tang) . n < 0, and -1 where |tang| <= 0.1.  four row-major, xn/xb
 * ns rows of 3. */
void orc_theta(int nx, int ny, int nz, int ns, const double *xi, const double *four, const double *xn,
               const double *xb, double *theta)
{
    long N = (long)nx * ny * nz;
    for (int isa = 0; isa < ns; isa++)
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++) {
                    long index = (long)i * ny * nz + (long)j * nz + k + isa * N;
                    long base = isa * N;
#define RE(ii, jj, kk) xi[2 * (base + (long)(ii) * ny * nz + (long)(jj) * nz + (kk))]
                    double gx = (RE((i + 1) % nx, j, k) - RE((i - 1 + nx) % nx, j, k)) / 2;
                    double gy = (RE(i, (j + 1) % ny, k) - RE(i, (j - 1 + ny) % ny, k)) / 2;
                    double gz = (RE(i, j, (k + 1) % nz) - RE(i, j, (k - 1 + nz) % nz)) / 2;
#undef RE
                    gx = gx * four[0] + gy * four[3] + gz * four[6];
                    gy = gx * four[1] + gy * four[4] + gz * four[7];
                    gz = gx * four[2] + gy * four[5] + gz * four[8];
                    const double *n = xn + 3 * isa, *b = xb + 3 * isa;
                    double t0 = gy * n[2] - gz * n[1], t1 = gz * n[0] - gx * n[2], t2 = gx * n[1] - gy * n[0];
                    double mag = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
                    if (mag <= 0.1) {
                        theta[index] = -1;
                    } else {
                        t0 /= mag;
                        t1 /= mag;
                        t2 /= mag;
                        theta[index] = acos(t0 * b[0] + t1 * b[1] + t2 * b[2]);
                        double dir = (b[1] * t2 - b[2] * t1) * n[0] + (b[2] * t0 - b[0] * t2) * n[1] +
                                     (b[0] * t1 - b[1] * t0) * n[2];
                        if (dir < 0.0) theta[index] = acos(-1) - theta[index];
                    }
                }
}

/* data_real_a = sum_b BB[r + a N + b N S] * data_fftw_b : internal_energy,
 * :1670-1695 (loop order isa, isb, grid). */
void orc_internal_energy(int nx, int ny, int nz, int ns, const double *BB,
                         const double *data_fftw, double *data_real)
{
    long N = (long)nx * ny * nz;
    memset(data_real, 0, sizeof(double) * 2 * N * ns);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (long r = 0; r < N; r++)
        for (int isa = 0; isa < ns; isa++)
            for (int isb = 0; isb < ns; isb++) {
                long index = r + isa * N, index2 = r + isb * N;
                long nb = r + isa * N + (long)isb * N * ns;
                data_real[2 * index] += data_fftw[2 * index2] * BB[nb];
                data_real[2 * index + 1] += data_fftw[2 * index2 + 1] * BB[nb];
            }
}

/* Explicit Euler update: update_order_parameter, :1746-1790, one rank.
 * xo may be NULL (no obstacles).  Reproduces quirk Q3: xirep/xiimp carry the
 * last unmasked cell's old value into masked cells' xinorm term.
 * out[0] = xinorm, out[1] = xiave. */
void orc_update(int nx, int ny, int nz, int ns, double *xi, const double *xo,
                const double *data_real, const double *dE, const double *tau, double cddt,
                double *out)
{
    long N = (long)nx * ny * nz;
    int nsize = nx * ny * nz;
    double xinormlocal = 0.0, xi_ave = 0.0, xirep = 0.0, xiimp = 0.0;
    for (int isa = 0; isa < ns; isa++)
        for (long r = 0; r < N; r++) {
            long index = r + isa * N;
            long na0 = 2 * index, na1 = na0 + 1;
            if (xo == NULL || xo[index] == 0.0) {
                xirep = xi[na0];
                xiimp = xi[na1];
                xi[na0] = xi[na0] - ((cddt) * (data_real[2 * index] / (nsize)-tau[isa] + dE[index]));
                xi[na1] = xi[na1] - ((cddt) * (data_real[2 * index + 1] / (nsize)));
            }
            xinormlocal += (xi[na0] - xirep) * (xi[na0] - xirep) + (xi[na1] - xiimp) * (xi[na1] - xiimp);
            xi_ave += xi[na0];
        }
    out[0] = sqrt(xinormlocal);
    out[1] = xi_ave;
}

/* ------------------------------------------------------------------------ */
/* One GL iteration as SolveGL::iterate runs it (solve_GL.cpp:86-101):       */
/* core_energy (on data_fftw = xi), forward FFTs, internal_energy, backward  */
/* FFTs, update.  work must hold 4*ns*N doubles; on return work[0:2nsN] is   */
/* data_fftw = FFT(xi_old) (the array the output step reads, quirk Q2).      */
/* out[0] xinorm, out[1] xiave, out[2] E_el, out[3] E_core (variational),    */
/* out[4] E_stress = sum_a tau_a sum_r Re xi_a, out[5] E_total,              */
/* out[6] E_core as the reference accumulates it (quirk Q9).                 */
/* All energies are those of the state before the update.                    */
/* ------------------------------------------------------------------------ */
void orc_gl_step(int nx, int ny, int nz, int ns, double *xi, const double *xo,
                 const double *BB, double cddt, const double *tau, int core_flag,
                 const double *core_par, int nm, const double *amp, double *work,
                 double *out)
{
    long N = (long)nx * ny * nz;
    double *data_fftw = work;
    double *data_real = work + 2 * N * ns;
    double *dE = (double *)malloc(sizeof(double) * N * ns);
    double e_var = 0.0;
    double e_ref = orc_core_energy(nx, ny, nz, ns, xi, core_flag, core_par, nm, amp, dE, &e_var);
    memcpy(data_fftw, xi, sizeof(double) * 2 * N * ns);
    for (int a = 0; a < ns; a++) orc_fft3d(nx, ny, nz, data_fftw + 2 * a * N, -1);
    orc_internal_energy(nx, ny, nz, ns, BB, data_fftw, data_real);
    /* elastic energy (1/2N) sum_k sum_a Re(conj(xihat_a) ghat_a), Parseval */
    double e_el = 0.0;
    for (long q = 0; q < N * ns; q++)
        e_el += data_fftw[2 * q] * data_real[2 * q] + data_fftw[2 * q + 1] * data_real[2 * q + 1];
    e_el /= 2.0 * (double)N;
    double e_st = 0.0;
    for (int a = 0; a < ns; a++) {
        double s = 0.0;
        for (long r = 0; r < N; r++) s += xi[2 * (r + a * N)];
        e_st += tau[a] * s;
    }
    for (int a = 0; a < ns; a++) orc_fft3d(nx, ny, nz, data_real + 2 * a * N, +1);
    double u[2];
    orc_update(nx, ny, nz, ns, xi, xo, data_real, dE, tau, cddt, u);
    free(dE);
    out[0] = u[0];
    out[1] = u[1];
    out[2] = e_el;
    out[3] = e_var;
    out[4] = e_st;
    out[5] = e_el + e_var - e_st;
    out[6] = e_ref;
}

/* ------------------------------------------------------------------------ */
/* Output-step diagnostics with both diag_style strain and stress, in the    */
/* order Output::compute runs them (output.cpp:133-188):                     */
/*  DiagStrain::compute (diag_strain.cpp:44-61): reset, strain(),            */
/*    total_average_strain();                                                */
/*  DiagStress::compute (diag_stress.cpp:43-73), comp_strain = 1 branch:     */
/*    reset local_sigma/ave_sigma, stress().                                 */
/* strain()  :3370-3469  (data_fftw = xihat of the pre-update xi, quirk Q2)  */
/* stressfree_strain() :3252-3326, total_average_strain() :3331-3365,        */
/* stress()  :3474-3553  (xi = post-update field).                           */
/* Outputs ave_eps[9], ave_sigma[9] (row-major 3x3); optional fields         */
/* data_eps / data_sigma (2*9*N each, layout of the reference) may be NULL.  */
/* mu_s is material->mu after Bmatrix; S11,S12,S44 are material compliances. */
/* ------------------------------------------------------------------------ */
void orc_diagnostics(int nx, int ny, int nz, int ns, const double *data_fftw,
                     const double *xi, const double *FF, const double *eps,
                     double C44, double C12, double C11, double S12, double S44,
                     double mu_s, const double *sigma_rot, double *ave_eps,
                     double *ave_sigma, double *data_eps_out, double *data_sigma_out)
{
    long N = (long)nx * ny * nz;
    int nsize = nx * ny * nz;
    double *data_strain = (double *)calloc(2 * 9 * N, sizeof(double));
    double *data_eps = (double *)calloc(2 * 9 * N, sizeof(double));
    double avepsd[9] = {0}, ave_epsd[9] = {0}, avepst[9] = {0}, avepsts[9] = {0};

    /* strain(): total strain in k space, :3407-3425 */
    for (int is = 0; is < ns; is++)
        for (int ida = 0; ida < 3; ida++)
            for (int idb = 0; idb < 3; idb++)
                for (long r = 0; r < N; r++) {
                    long index = r + is * N;
                    long index2 = r + ida * N + idb * N * 3;
                    long nb = r + is * N + (long)ida * N * ns + (long)idb * N * ns * 3;
                    data_strain[2 * index2] += data_fftw[2 * index] * FF[nb];
                    data_strain[2 * index2 + 1] += data_fftw[2 * index + 1] * FF[nb];
                }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            orc_fft3d(nx, ny, nz, data_strain + 2 * (i * N + j * N * 3), +1);
    for (long q = 0; q < 2 * 9 * N; q++) data_strain[q] = data_strain[q] / (nsize);

    /* stressfree_strain(): :3252-3326 */
    {
        double S[81];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++)
                        S[27 * i + 9 * j + 3 * k + l] =
                            S44 / 4 * (DELTA(i, k) * DELTA(j, l) + DELTA(i, l) * DELTA(j, k)) +
                            S12 * DELTA(i, j) * DELTA(k, l);
        for (int is = 0; is < ns; is++)
            for (int ida = 0; ida < 3; ida++)
                for (int idb = 0; idb < 3; idb++) {
                    for (long r = 0; r < N; r++)
                        avepsd[3 * ida + idb] += eps[9 * is + 3 * ida + idb] * xi[2 * (r + is * N)];
                    ave_epsd[3 * ida + idb] = avepsd[3 * ida + idb];
                    ave_epsd[3 * ida + idb] = ave_epsd[3 * ida + idb] / nsize;
                }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        avepst[3 * i + j] += S[27 * i + 9 * j + 3 * k + l] * sigma_rot[3 * k + l] * mu_s;
                        avepsts[3 * i + j] += S[27 * i + 9 * j + 3 * k + l] * sigma_rot[3 * k + l] * mu_s;
                    }
                avepst[3 * i + j] += ave_epsd[3 * i + j];
            }
    }
    /* strain() tail: add avepst, symmetrise, :3437-3464 */
    for (int ida = 0; ida < 3; ida++)
        for (int idb = 0; idb < 3; idb++)
            for (long r = 0; r < N; r++)
                data_strain[2 * (r + ida * N + idb * N * 3)] += avepst[3 * ida + idb];
    for (int ida = 0; ida < 3; ida++)
        for (int idb = 0; idb < 3; idb++)
            for (long r = 0; r < N; r++) {
                long index = r + ida * N + idb * N * 3;
                long index2 = r + idb * N + ida * N * 3;
                data_eps[2 * index] = (data_strain[2 * index] + data_strain[2 * index2]) / 2.0;
                data_eps[2 * index + 1] = (data_strain[2 * index + 1] + data_strain[2 * index2 + 1]) / 2.0;
            }
    /* total_average_strain(): :3331-3365 */
    for (int ida = 0; ida < 3; ida++)
        for (int idb = 0; idb < 3; idb++) {
            double s = 0.0;
            for (long r = 0; r < N; r++) s += data_eps[2 * (r + ida * N + idb * N * 3)];
            ave_eps[3 * ida + idb] = s / nsize;
        }

    /* stress(): :3474-3553 */
    {
        double C[81];
        orc_stiffness(C44, C12, C);
        (void)C11;
        double *data_epsd = (double *)calloc(2 * 9 * N, sizeof(double));
        double *data_sigma = data_sigma_out ? data_sigma_out : (double *)malloc(sizeof(double) * 2 * 9 * N);
        memset(data_sigma, 0, sizeof(double) * 2 * 9 * N);
        for (int is = 0; is < ns; is++)
            for (int ida = 0; ida < 3; ida++)
                for (int idb = 0; idb < 3; idb++)
                    for (long r = 0; r < N; r++)
                        data_epsd[2 * (r + ida * N + idb * N * 3)] +=
                            eps[9 * is + 3 * ida + idb] * xi[2 * (r + is * N)];
        for (int ida = 0; ida < 3; ida++)
            for (int idb = 0; idb < 3; idb++) {
                double local = 0.0;
                for (long r = 0; r < N; r++) {
                    long na = 2 * (r + ida * N + idb * N * 3);
                    for (int m = 0; m < 3; m++)
                        for (int l = 0; l < 3; l++) {
                            long na0 = 2 * (r + m * N + l * N * 3);
                            data_sigma[na] += C[27 * ida + 9 * idb + 3 * m + l] *
                                              (data_strain[na0] - data_epsd[na0] - avepsts[3 * m + l]);
                            data_sigma[na + 1] = 0.0;
                        }
                    data_sigma[na] += sigma_rot[3 * ida + idb] * mu_s;
                    local += data_sigma[na];
                }
                ave_sigma[3 * ida + idb] = local / nsize;
            }
        free(data_epsd);
        if (!data_sigma_out) free(data_sigma);
    }
    if (data_eps_out) memcpy(data_eps_out, data_eps, sizeof(double) * 2 * 9 * N);
    free(data_strain);
    free(data_eps);
}

/* ------------------------------------------------------------------------ */
/* RANMAR uniform generator: RanMars::init / uniform, random_mars.cpp:41-101.*/
double orc_ranmar_uniform(double *st);
/* state: 97 u values + c, cd, cm, i97, j97 packed in a double[103].         */
/* ------------------------------------------------------------------------ */
void orc_ranmar_init(int seed, double *st)
{
    int ij, kl, i, j, k, l, ii, jj, m;
    double s, t;
    while (seed > 900000000) seed -= 900000000;
    ij = (seed - 1) / 30082;
    kl = (seed - 1) - 30082 * ij;
    i = (ij / 177) % 177 + 2;
    j = ij % 177 + 2;
    k = (kl / 169) % 178 + 1;
    l = kl % 169;
    for (ii = 1; ii <= 97; ii++) {
        s = 0.0;
        t = 0.5;
        for (jj = 1; jj <= 24; jj++) {
            m = ((i * j) % 179) * k % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s = s + t;
            t = 0.5 * t;
        }
        st[ii] = s;
    }
    st[98] = 362436.0 / 16777216.0;
    st[99] = 7654321.0 / 16777216.0;
    st[100] = 16777213.0 / 16777216.0;
    st[101] = 97;
    st[102] = 33;
    orc_ranmar_uniform(st); /* RanMars::init discards one value */
}

double orc_ranmar_uniform(double *st)
{
    int i97 = (int)st[101], j97 = (int)st[102];
    double uni = st[i97] - st[j97];
    if (uni < 0.0) uni += 1.0;
    st[i97] = uni;
    i97--;
    if (i97 == 0) i97 = 97;
    j97--;
    if (j97 == 0) j97 = 97;
    st[101] = i97;
    st[102] = j97;
    st[98] -= st[99];
    if (st[98] < 0.0) st[98] += st[100];
    uni -= st[98];
    if (uni < 0.0) uni += 1.0;
    return uni;
}

/* n successive RANMAR uniforms into out (same stream as orc_ranmar_uniform). */
void orc_ranmar_fill(double *st, long n, double *out)
{
    for (long q = 0; q < n; q++) out[q] = orc_ranmar_uniform(st);
}
