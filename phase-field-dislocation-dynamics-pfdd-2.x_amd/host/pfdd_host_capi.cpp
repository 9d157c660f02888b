// pfdd_host_capi.cpp -- C ABI of the host setup helpers (include/pfdd_host.h).
#include "../../include/pfdd_host.h"

#include <cstring>
#include <string>
#include <vector>

#include "pfdd_host.hpp"

using namespace pfdd_host;

namespace {

void geometry_from(const double *prim9, const double *scale3, Geometry &g)
{
    if (prim9) g.set_primitive(prim9, prim9 + 3, prim9 + 6);
    if (scale3) g.set_scale(scale3[0], scale3[1], scale3[2]);
    else g.set_scale(1.0, 1.0, 1.0);
}

} // namespace

extern "C" {

int pfdd_host_material_preset(const char *preset, pfdd_host_material *m)
{
    if (!m) return 1;
    std::memset(m, 0, sizeof(*m));
    Material mat;
    if (preset && *preset && !mat.preset(preset)) return 1;
    m->a = mat.a;
    m->b = mat.b;
    m->mu = mat.mu;
    m->young = mat.young;
    m->isf = mat.isf;
    m->usf = mat.usf;
    m->An = mat.An;
    m->Cn = mat.Cn;
    m->dslip = 0.8165;
    return 0;
}

int pfdd_host_material_derive(pfdd_host_material *m)
{
    if (!m || !(m->mu > 0) || !(m->young > 0)) return 1;
    Material mat;
    mat.mu = m->mu;
    mat.young = m->young;
    mat.derive();
    m->dslip = mat.dslip;
    m->C44 = mat.C44;
    m->C12 = mat.C12;
    m->C11 = mat.C11;
    m->nu = mat.nu;
    m->S11 = mat.S11;
    m->S12 = mat.S12;
    m->S44 = mat.S44;
    m->mu_s = mat.mu_s;
    m->xnu_s = mat.xnu_s;
    return 0;
}

int pfdd_host_slip(const char *app, const double *prim9, const double *scale3, int *ns, double *xn,
                   double *xb, double *four9, double *cd, double *timestep)
{
    if (!app || !ns || !xn || !xb) return 1;
    SlipTable t;
    if (!slip_table(app, t)) return 1;
    Geometry g;
    geometry_from(prim9, scale3, g);
    rotate_slip(t, g);
    *ns = t.slip_systems;
    std::memcpy(xn, t.xn.data(), sizeof(double) * 3 * t.slip_systems);
    std::memcpy(xb, t.xb.data(), sizeof(double) * 3 * t.slip_systems);
    if (four9)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) four9[3 * r + c] = g.four[r][c];
    if (cd) *cd = t.CD;
    if (timestep) *timestep = t.timestep;
    return 0;
}

int pfdd_host_resolss_non_schmid(const char *app, const double *prim9, const double *scale3,
                                 const double *sigma9, double dslip, double angle, double w1, double w2, double w3,
                                 double *tau)
{
    if (!app || !sigma9 || !tau) return 1;
    SlipTable t;
    if (!slip_table(app, t)) return 1;
    Geometry g;
    geometry_from(prim9, scale3, g);
    rotate_slip(t, g);
    double s[3][3], sr[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s[r][c] = sigma9[3 * r + c];
    rotate_stress(g, s, sr);
    resolss_non_schmid(t, sr, dslip, angle, w1, w2, w3, tau);
    return 0;
}

int pfdd_host_resolss(const char *app, const double *prim9, const double *scale3, const double *sigma9,
                      double dslip, double *tau, double *sigma_rot9)
{
    if (!app || !sigma9 || !tau) return 1;
    SlipTable t;
    if (!slip_table(app, t)) return 1;
    Geometry g;
    geometry_from(prim9, scale3, g);
    rotate_slip(t, g);
    double s[3][3], sr[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s[r][c] = sigma9[3 * r + c];
    rotate_stress(g, s, sr);
    resolss(t, sr, dslip, tau);
    if (sigma_rot9)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) sigma_rot9[3 * r + c] = sr[r][c];
    return 0;
}

int pfdd_host_core_coef(const pfdd_host_material *m, int flag, double *coef, int *ncoef)
{
    if (!m || !coef) return 1;
    Material mat;
    mat.b = m->b;
    mat.isf = m->isf;
    mat.usf = m->usf;
    mat.An = m->An;
    mat.mu_s = m->mu_s;
    mat.dslip = m->dslip;
    std::vector<double> c = core_coefficients(mat, flag);
    if (c.empty()) return 1;
    for (size_t q = 0; q < c.size(); ++q) coef[q] = c[q];
    if (ncoef) *ncoef = (int)c.size();
    return 0;
}

int pfdd_host_ranmar_fill(int seed, long long n, double *out)
{
    if (!out || n < 0 || seed <= 0) return 1;
    RanMars r(seed);
    for (long long q = 0; q < n; ++q) out[q] = r.uniform();
    return 0;
}

int pfdd_host_init_random_loops(int nx, int ny, int nz, const char *app, int nloops, double rmin,
                                double rmax, double noise, int seed, double *xi, double *loops)
{
    if (!app || !xi || nx <= 0 || ny <= 0 || nz <= 0 || seed <= 0) return 1;
    SlipTable t;
    if (!slip_table(app, t)) return 1;
    Grid g{nx, ny, nz, t.slip_systems};
    std::vector<double> lo;
    init_random_loops(g, t, nloops, rmin, rmax, noise, seed, xi, &lo);
    if (loops) std::memcpy(loops, lo.data(), sizeof(double) * lo.size());
    return 0;
}

int pfdd_host_init_sxtal(const char *style, int nx, int ny, int nz, int ns, double *xi, double *xo)
{
    if (!style || !xi || nx <= 0 || ny <= 0 || nz <= 0 || ns <= 0) return 1;
    Grid g{nx, ny, nz, ns};
    const std::string s = style;
    if (s == "loop") init_sxtal_loop(g, xi);
    else if (s == "frank_read") init_sxtal_frank_read(g, xi, xo);
    else if (s == "edge") init_sxtal_edge(g, xi);
    else if (s == "3SBCC") init_sxtal_3sbcc(g, xi);
    else if (s == "StraightNonOrtho") init_sxtal_straight_nonortho(g, xi);
    else if (s == "StraightOrtho") init_sxtal_straight_ortho(g, xi);
    else if (init_sxtal_more(s, g, Geometry(), 1.0, (double)nx, xi, xo)) return 0;
    else if (s == "1LOrthoBCC") init_sxtal_1lorthobcc(g, 1.0, 1.0, xi);
    else return 1;
    if (xo && s != "frank_read")
        for (long long q = 0; q < (long long)ns * g.N(); ++q) xo[q] = 0.0;
    return 0;
}

} // extern "C"

static void cond_params(const double *p, Material &m)
{
    m.condsigma = p[0], m.mfpel = p[1], m.alpha_resistivity = p[2], m.alpha_mfpel = p[3], m.Rgrain = p[4];
    m.beta_disloc = p[5], m.rho_disloc = p[6], m.dlayer = p[7], m.Nlayers = (int)p[8], m.Ngrains = (int)p[9];
    m.pinterface = p[10];
}

extern "C" int pfdd_host_cond_preset(const char *preset, double *par)
{
    if (!par) return 1;
    Material m;
    if (preset && *preset && !m.preset(preset)) return 1;
    const double v[11] = {m.condsigma, m.mfpel, m.alpha_resistivity, m.alpha_mfpel, m.Rgrain, m.beta_disloc,
                          m.rho_disloc, m.dlayer, (double)m.Nlayers, (double)m.Ngrains, m.pinterface};
    for (int q = 0; q < 11; ++q) par[q] = v[q];
    return 0;
}

extern "C" int pfdd_host_cond_layers(const double *parA, const double *parB, int ny, double temperature,
                                     double *sigma, double *dsigma)
{
    if (!parA || !parB || !sigma || !dsigma) return 1;
    Material A, B;
    cond_params(parA, A);
    cond_params(parB, B);
    std::vector<double> s, d;
    if (!cond_layers(A, B, ny, temperature, s, d)) return 1;
    for (int j = 0; j < ny; ++j) sigma[j] = s[j], dsigma[j] = d[j];
    return 0;
}

// ------------------------------------------- raw-array forms for the plugin
extern "C" int pfdd_host_rotate_stress_raw(const double *prim9, const double *invprim9, double norvol,
                                           const double *sigma9, double *sigma_rot9)
{
    if (!prim9 || !invprim9 || !sigma9 || !sigma_rot9) return 1;
    Geometry g;
    double s[3][3], sr[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            g.prim[i][j] = prim9[3 * i + j];
            g.invprim[i][j] = invprim9[3 * i + j];
            s[i][j] = sigma9[3 * i + j];
        }
    g.norvol = norvol;
    rotate_stress(g, s, sr);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sigma_rot9[3 * i + j] = sr[i][j];
    return 0;
}

static void table_from(int ns, const double *xn, const double *xb, SlipTable &t)
{
    t.slip_systems = ns;
    t.xn.assign(xn ? xn : xb, (xn ? xn : xb) + 3 * ns);
    t.xb.assign(xb, xb + 3 * ns);
}

extern "C" int pfdd_host_resolss_raw(int ns, const double *xn, const double *xb, const double *sigma_rot9,
                                     double dslip, double *tau)
{
    if (ns < 1 || !xn || !xb || !sigma_rot9 || !tau || dslip == 0.0) return 1;
    SlipTable t;
    table_from(ns, xn, xb, t);
    double sr[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sr[i][j] = sigma_rot9[3 * i + j];
    resolss(t, sr, dslip, tau);
    return 0;
}

extern "C" int pfdd_host_resolss_non_schmid_raw(int ns, const double *xb, const double *sigma_rot9, double dslip,
                                                double angle, double w1, double w2, double w3, double *tau)
{
    if (ns < 1 || !xb || !sigma_rot9 || !tau || dslip == 0.0) return 1;
    SlipTable t;
    table_from(ns, nullptr, xb, t);
    double sr[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sr[i][j] = sigma_rot9[3 * i + j];
    resolss_non_schmid(t, sr, dslip, angle, w1, w2, w3, tau);
    return 0;
}

extern "C" double pfdd_host_usfe_angle_tau_raw(const pfdd_host_material *m, const double *par5,
                                               const double *sigma_rot9, double angle)
{
    if (!m || !par5 || !sigma_rot9) return 0.0;
    Material mat;
    mat.mu_s = m->mu_s;
    mat.dslip = m->dslip;
    mat.b = m->b;
    mat.a_slope = par5[0], mat.a_b = par5[1], mat.b_b = par5[2], mat.c_slope = par5[3], mat.c_b = par5[4];
    double sr[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) sr[i][j] = sigma_rot9[3 * i + j];
    return usfe_angle_tau(mat, sr, angle);
}

extern "C" int pfdd_host_project_core_energy(int nx, int ny, int nz, int ns, const double *xb, const double *xi,
                                             double *delta, double *ddelta)
{
    if (nx < 1 || ny < 1 || nz < 1 || ns < 1 || !xb || !xi || !delta || !ddelta) return 1;
    project_core_energy(nx, ny, nz, ns, xb, xi, delta, ddelta);
    return 0;
}
