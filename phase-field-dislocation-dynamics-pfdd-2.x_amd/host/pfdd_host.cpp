// pfdd_host.cpp -- see pfdd_host.hpp.
#include "pfdd_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace pfdd_host {

// ------------------------------------------------------------------ Material
bool Material::preset(const std::string &name)
{
    // set_constants, material.cpp:176-276 (elastic and GSF part of the presets)
    if (name == "Ni") {
        a = 3.52E-10;
        b = a / std::sqrt(2.0);
    } else if (name == "Cu") {
        a = 3.64E-10;
        b = a / std::sqrt(2.0);
        mu = 54.5E9;
        young = 144.4E9;
        c0 = 344.948E-3;
        c1 = -85.399E-3;
        c2 = -40.421E-3;
        c3 = 12.396E-3;
        c4 = -0.674E-3;
        a1 = -131.393E-3;
        a3 = -19.816E-3;
        isf = 38.5E-3;
        usf = 163.7E-3;
    } else if (name == "Pd") {
        a = 3.950E-10;
        b = a / std::sqrt(2.0);
        mu = 45.7E9;
        young = 134.60E9;
        usf = 0.76308;
        An = usf;
    } else {
        return false;
    }
    return true;
}

void Material::derive()
{
    // material.cpp:156-167
    C44 = mu;
    nu = young / 2.0 / (mu)-1.0;
    C12 = 2.0 * nu * C44 / (1.0 - 2.0 * nu);
    C11 = 2.0 * C44 + C12;
    ll = C12;
    S11 = 1.0 / (young);
    S12 = -nu / (young);
    S44 = 2 * (S11 - S12);
    dslip = 0.8165;
    // Bmatrix's Voigt average (fft_fftw_slab.cpp:487-496)
    mu_s = C44 - (2.0 * C44 + C12 - C11) / 5.0;
    ll_s = C12 - (2.0 * C44 + C12 - C11) / 5.0;
    young_s = mu_s * (3 * ll_s + 2 * mu_s) / (ll_s + mu_s);
    xnu_s = young_s / 2.0 / mu_s - 1.0;
}

// ------------------------------------------------------------------ Geometry
void Geometry::set_primitive(const double x[3], const double y[3], const double z[3])
{
    // fft.cpp:101-133: columns are the normalised x, y, z vectors
    const double *v[3] = {x, y, z};
    for (int c = 0; c < 3; ++c) {
        const double n = std::sqrt(v[c][0] * v[c][0] + v[c][1] * v[c][1] + v[c][2] * v[c][2]);
        for (int r = 0; r < 3; ++r) prim[r][c] = v[c][r] / n;
    }
    primitive = true;
    // fft.cpp:155-169
    const double(*p)[3] = prim;
    norvol = p[1][0] * p[2][1] * p[0][2] - p[2][0] * p[1][1] * p[0][2] + p[2][0] * p[0][1] * p[1][2] -
             p[0][0] * p[2][1] * p[1][2] + p[0][0] * p[1][1] * p[2][2] - p[1][0] * p[0][1] * p[2][2];
    invprim[0][0] = (p[1][1] * p[2][2] - p[2][1] * p[1][2]) / norvol;
    invprim[0][1] = (p[2][1] * p[0][2] - p[0][1] * p[2][2]) / norvol;
    invprim[0][2] = (p[0][1] * p[1][2] - p[1][1] * p[0][2]) / norvol;
    invprim[1][0] = (p[1][2] * p[2][0] - p[2][2] * p[1][0]) / norvol;
    invprim[1][1] = (p[2][2] * p[0][0] - p[0][2] * p[2][0]) / norvol;
    invprim[1][2] = (p[0][2] * p[1][0] - p[1][2] * p[0][0]) / norvol;
    invprim[2][0] = (p[1][0] * p[2][1] - p[2][0] * p[1][1]) / norvol;
    invprim[2][1] = (p[2][0] * p[0][1] - p[0][0] * p[2][1]) / norvol;
    invprim[2][2] = (p[0][0] * p[1][1] - p[1][0] * p[0][1]) / norvol;
}

void Geometry::set_scale(double sa, double sb, double sc)
{
    // App::set_scale, app.cpp:383-431
    for (int r = 0; r < 3; ++r) {
        sclprim[r][0] = sa * prim[r][0];
        sclprim[r][1] = sb * prim[r][1];
        sclprim[r][2] = sc * prim[r][2];
    }
    const double(*s)[3] = sclprim;
    vol = s[1][0] * s[2][1] * s[0][2] - s[2][0] * s[1][1] * s[0][2] + s[2][0] * s[0][1] * s[1][2] -
          s[0][0] * s[2][1] * s[1][2] + s[0][0] * s[1][1] * s[2][2] - s[1][0] * s[0][1] * s[2][2];
    four[0][0] = (s[1][1] * s[2][2] - s[2][1] * s[1][2]) / vol;
    four[0][1] = (s[2][1] * s[0][2] - s[0][1] * s[2][2]) / vol;
    four[0][2] = (s[0][1] * s[1][2] - s[1][1] * s[0][2]) / vol;
    four[1][0] = (s[1][2] * s[2][0] - s[2][2] * s[1][0]) / vol;
    four[1][1] = (s[2][2] * s[0][0] - s[0][2] * s[2][0]) / vol;
    four[1][2] = (s[0][2] * s[1][0] - s[1][2] * s[0][0]) / vol;
    four[2][0] = (s[1][0] * s[2][1] - s[2][0] * s[1][1]) / vol;
    four[2][1] = (s[2][0] * s[0][1] - s[0][0] * s[2][1]) / vol;
    four[2][2] = (s[0][0] * s[1][1] - s[1][0] * s[0][1]) / vol;
    for (int r = 0; r < 3; ++r)
        normfour[r] = std::sqrt(four[r][0] * four[r][0] + four[r][1] * four[r][1] + four[r][2] * four[r][2]);
}

// --------------------------------------------------------------- slip tables
bool slip_table(const std::string &app, SlipTable &t, double oflag)
{
    t = SlipTable();
    t.app = app;
    if (app == "fcc") {
        // AppFCC (app_fcc.cpp:29-44, 211-286)
        const double s3 = std::sqrt(3.0);
        const int sg[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};
        static const double b[12][3] = {
            {-0.5, 0.5, 0.0}, {0.0, -0.5, 0.5}, {0.5, 0.0, -0.5}, {-0.5, -0.5, 0.0},
            {0.5, 0.0, 0.5},  {0.0, -0.5, 0.5}, {-0.5, -0.5, 0.0}, {0.5, 0.0, -0.5},
            {0.0, -0.5, -0.5}, {-0.5, 0.5, 0.0}, {0.5, 0.0, 0.5},  {0.0, -0.5, -0.5}};
        t.slip_systems = 12;
        t.num_planes = 4;
        t.CD = 0.5;
        t.timestep = 1.0;
        t.default_core = 1;
        for (int a = 0; a < 12; ++a)
            for (int c = 0; c < 3; ++c) {
                t.xn.push_back(sg[a / 3][c] / s3);
                t.xb.push_back(b[a][c]);
            }
        return true;
    }
    if (app == "3d-hcp-basal") {
        // App3DhcpBasal (app_3d_hcp_basal.cpp:28-45, 216-246): b at 0, +60, -60 deg
        t.slip_systems = 3;
        t.num_planes = 3;
        t.CD = 0.5;
        t.default_core = 1;
        const double ang[3] = {oflag * M_PI / 2.0, oflag * M_PI / 2.0 + M_PI / 3.0,
                               oflag * M_PI / 2.0 - M_PI / 3.0};
        for (int a = 0; a < 3; ++a) {
            t.xn.insert(t.xn.end(), {0.0, 0.0, 1.0});
            t.xb.insert(t.xb.end(), {std::cos(ang[a]), std::sin(ang[a]), 0.0});
        }
        return true;
    }
    if (app == "basal3") {
        // three unit Burgers vectors at 0, 120, 240 deg on the plane n = z
        t.slip_systems = 3;
        t.num_planes = 3;
        t.CD = 0.5;
        t.default_core = 3;
        for (int a = 0; a < 3; ++a) {
            const double th = 2.0 * M_PI * a / 3.0;
            t.xn.insert(t.xn.end(), {0.0, 0.0, 1.0});
            t.xb.insert(t.xb.end(), {std::cos(th), std::sin(th), 0.0});
        }
        return true;
    }
    if (app == "frank_read") {
        // AppFrankRead (app_frank_read.cpp:29-50, 96-120)
        t.slip_systems = 2;
        t.num_planes = 2;
        t.CD = 0.35;
        t.default_core = 1;
        t.rotate = false; // AppFrankRead::set_slip has no rotation block
        t.xn = {0.0, 0.7071, 0.7071, 0.7071, 0, 0.7071};
        t.xb = {0.57735, 0.57735, -0.57735, 0.57735, 0.57735, -0.57735};
        return true;
    }
    return false;
}

void rotate_slip(SlipTable &t, const Geometry &g)
{
    if (!t.rotate) return;
    for (int i = 0; i < t.slip_systems; ++i) {
        double *v[2] = {&t.xn[3 * i], &t.xb[3 * i]};
        for (double *w : v) {
            const double a = (w[0] * g.four[0][0] + w[1] * g.four[0][1] + w[2] * g.four[0][2]) / g.normfour[0];
            const double b = (w[0] * g.four[1][0] + w[1] * g.four[1][1] + w[2] * g.four[1][2]) / g.normfour[1];
            const double c = (w[0] * g.four[2][0] + w[1] * g.four[2][1] + w[2] * g.four[2][2]) / g.normfour[2];
            w[0] = a;
            w[1] = b;
            w[2] = c;
        }
    }
}

void rotate_stress(const Geometry &g, const double sigma[3][3], double sigma_rot[3][3])
{
    const bool ortho = std::fabs(g.norvol) < 1.01 && std::fabs(g.norvol) > 0.99;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            if (ortho) {
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) s += g.prim[i][k] * sigma[k][l] * g.invprim[l][j];
            } else {
                s = sigma[i][j];
            }
            sigma_rot[i][j] = s;
        }
}

void resolss(const SlipTable &t, const double sigma_rot[3][3], double dslip, double *tau)
{
    for (int is = 0; is < t.slip_systems; ++is) {
        double v = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) v = v + (sigma_rot[i][j] * t.xn[3 * is + j] * t.xb[3 * is + i]) / dslip;
        tau[is] = v;
    }
}

std::vector<double> core_coefficients(const Material &m, int flag)
{
    const double mdb = m.mu_s * m.dslip * m.b;
    if (flag == 1) return {m.An / mdb};
    if (flag == 3) {
        const double isf = m.isf / mdb;
        const double Cn = (m.usf - (isf / 2.0)) / mdb;
        return {isf, Cn};
    }
    return {};
}

// -------------------------------------------------------------------- RANMAR
RanMars::RanMars(int seed)
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
        u[ii] = s;
    }
    c = 362436.0 / 16777216.0;
    cd = 7654321.0 / 16777216.0;
    cm = 16777213.0 / 16777216.0;
    i97 = 97;
    j97 = 33;
    uniform();
}

double RanMars::uniform()
{
    double uni = u[i97] - u[j97];
    if (uni < 0.0) uni += 1.0;
    u[i97] = uni;
    if (--i97 == 0) i97 = 97;
    if (--j97 == 0) j97 = 97;
    c -= cd;
    if (c < 0.0) c += cm;
    uni -= c;
    if (uni < 0.0) uni += 1.0;
    return uni;
}

// ------------------------------------------------------------ initial fields
static void miller(const SlipTable &t, int a, int hkl[3])
{
    double mn = 1e300;
    for (int c = 0; c < 3; ++c)
        if (std::fabs(t.xn[3 * a + c]) > 1e-6) mn = std::fmin(mn, std::fabs(t.xn[3 * a + c]));
    for (int c = 0; c < 3; ++c) hkl[c] = (int)std::lround(t.xn[3 * a + c] / mn);
}

static inline int wrapd(int d, int n)
{
    d %= n;
    if (d < -n / 2) d += n;
    if (d >= n - n / 2) d -= n;
    return d;
}

void init_random_loops(const Grid &g, const SlipTable &t, int nloops, double rmin, double rmax,
                       double noise, int seed, double *xi, std::vector<double> *loops_out)
{
    const long long N = g.N();
    RanMars rng(seed);
    for (long long q = 0; q < (long long)g.S * N; ++q) {
        xi[2 * q] = noise > 0 ? 2.0 * noise * (rng.uniform() - 0.5) : 0.0;
        xi[2 * q + 1] = 0.0;
    }
    const int nplanes = t.num_planes > 0 ? t.num_planes : 1;
    const int per_plane = g.S / nplanes > 0 ? g.S / nplanes : 1;
    for (int L = 0; L < nloops; ++L) {
        const int p = std::min(nplanes - 1, (int)(rng.uniform() * nplanes));
        const int s = std::min(per_plane - 1, (int)(rng.uniform() * per_plane));
        const int a = std::min(g.S - 1, p * per_plane + s);
        const int ci = std::min(g.nx - 1, (int)(rng.uniform() * g.nx));
        const int cj = std::min(g.ny - 1, (int)(rng.uniform() * g.ny));
        const int ck = std::min(g.nz - 1, (int)(rng.uniform() * g.nz));
        const double R = rmin + (rmax - rmin) * rng.uniform();
        if (loops_out) loops_out->insert(loops_out->end(), {(double)a, (double)ci, (double)cj, (double)ck, R});
        int hkl[3];
        miller(t, a, hkl);
        const int r = (int)std::ceil(R);
        for (int di = -r; di <= r; ++di)
            for (int dj = -r; dj <= r; ++dj)
                for (int dk = -r; dk <= r; ++dk) {
                    if (hkl[0] * di + hkl[1] * dj + hkl[2] * dk != 0) continue;
                    if (di * di + dj * dj + dk * dk > R * R) continue;
                    const int i = ((ci + di) % g.nx + g.nx) % g.nx;
                    const int j = ((cj + dj) % g.ny + g.ny) % g.ny;
                    const int k = ((ck + dk) % g.nz + g.nz) % g.nz;
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)a * N;
                    xi[2 * idx] += 1.0;
                }
    }
    (void)wrapd;
}

void init_sxtal_loop(const Grid &g, double *xi)
{
    const long long N = g.N();
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const int d0 = -1 * (j - g.ny / 2) + (k - g.nz / 2);
                    const int d1 = (j - g.ny / 2) - (k - g.nz / 2);
                    const int d2 = (j - g.ny / 2) + (k - g.nz / 2);
                    xi[2 * idx] = (is == 0 && i == 0 && d0 * d0 + d1 * d1 + d2 * d2 < 3 * 16 * 16) ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                }
}

void init_sxtal_frank_read(const Grid &g, double *xi, double *xo)
{
    const long long N = g.N();
    const int fr_halflen = 10;
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const bool on = is == 1 && i < g.nx / 2 + fr_halflen && i >= g.nx / 2 - fr_halflen &&
                                    j <= g.ny / 2 && k == g.nz / 2;
                    xi[2 * idx] = on ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                    if (xo) xo[idx] = (j != 0 && j != g.ny / 2) ? 1.0 : 0.0;
                }
}

void init_sxtal_edge(const Grid &g, double *xi)
{
    const long long N = g.N();
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const bool on = is == 0 && i == 0 &&
                                    ((k <= 3 * j && (k + g.nz / 2 > 3 * j)) || (k - g.nz / 2) > 3 * j);
                    xi[2 * idx] = on ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                }
}

void init_sxtal_3sbcc(const Grid &g, double *xi)
{
    const long long N = g.N();
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const bool on = is == 0 && i == 0 && k >= g.nz / 4 && k < g.nz * 3 / 4;
                    xi[2 * idx] = on ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                }
}

} // namespace pfdd_host
