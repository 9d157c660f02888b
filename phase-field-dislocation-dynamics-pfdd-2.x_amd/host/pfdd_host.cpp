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
    // set_constants, material.cpp:176-835: elastic constants, Burgers vector
    // rule, sine-core (isf, usf, An), extended gamma-surface (c0..c4, a1, a3)
    // and pyramidal-II Fourier (aa0..aa4, bb1..bb4) constants of each named
    // preset, and the conduction constants of those that have them.
    // b: 0 = a/sqrt2 (fcc), 1 = sqrt3 a/2 (bcc), 2 = given, 3 = not set
    struct Row {
        const char *name;
        double a;
        int brule;
        double b, mu, young;
        double c[7]; // c0..c4, a1, a3
        double isf, usf, An;
        double pyr[9]; // aa0..aa4, bb1..bb4
    };
    static const double MG_PYR[9] = {216.6E-3, -82.84E-3, -116.8E-3, -5.873E-3, -11.27E-3,
                                     -51.94E-3, 22.04E-3, 9.44E-3, -0.2501E-3};
    static const Row rows[] = {
        {"Ni", 3.52E-10, 0, 0, 99.7E9, 255.7E9,
         {503.144E-3, -106.533E-3, -83.741E-3, 28.777E-3, -2.916E-3, -123.917E-3, -38.533E-3}, 144.5E-3, 289.0E-3, 0, {}},
        {"Al", 4.04E-10, 0, 0, 28.4E9, 76.0E9,
         {149.56346E-3, -4.23464E-3, -74.39331E-3, 33.84682E-3, -2.39438E-3, 67.75268E-3, -30.88995E-3}, 140.2E-3,
         177.0E-3, 0, {}},
        {"Au", 4.17E-10, 0, 0, 18.8E9, 54.0E9,
         {161.832E-3, -40.659E-3, -17.320E-3, 3.197E-3, 0.472E-3, -64.718E-3, -13.282E-3}, 27.9E-3, 66.5E-3, 0, {}},
        {"Cu", 3.64E-10, 0, 0, 54.5E9, 144.4E9,
         {344.948E-3, -85.399E-3, -40.421E-3, 12.396E-3, -0.674E-3, -131.393E-3, -19.816E-3}, 38.5E-3, 163.7E-3, 0, {}},
        {"Pd", 3.950E-10, 0, 0, 45.7E9, 125.8E9,
         {308.65990E-3, -63.66363E-3, -48.76003E-3, 12.024463E-3, -1.11869E-3, -61.34161E-3, -23.33224E-3}, 138.1E-3,
         197.9E-3, 0, {}},
        {"Ag", 4.16E-10, 0, 0, 27.5E9, 74.8E9,
         {179.573E-3, -32.837E-3, -39.578E-3, 14.700E-3, -0.989E-3, -43.791E-3, -17.254E-3}, 17.8E-3, 100.4E-3, 0, {}},
        {"Pt", 3.98E-10, 0, 0, 53.2E9, 149.0E9,
         {223.27233E-3, -15.38077E-3, -75.36747E-3, 21.70531E-3, -2.52876E-3, 58.60391E-3, -43.19019E-3}, 253.7E-3,
         257.5E-3, 0, {}},
        {"Rh", 3.84E-10, 0, 0, 154.1E9, 384.9E9,
         {556.35590E-3, -46.42878E-3, -183.45940E-3, 55.25737E-3, -5.04421E-3, 1.44561E-3, -72.22364E-3}, 184.6E-3,
         454.1E-3, 0, {}},
        {"Ir", 3.88E-10, 0, 0, 222.3E9, 549.7E9,
         {617.61574E-3, 3.28767E-3, -283.1158E-3, 96.12464E-3, -10.56298E-3, 157.20734E-3, -109.4680E-3}, 324.4E-3,
         614.9E-3, 0, {}},
        {"Mg", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {112.6205E-3, 0.7200E-3, -58.3781E-3, 28.1313E-3, -3.9212E-3, 25.2325E-3, -23.8545E-3}, 29.3216E-3,
         88.4906E-3, 0, {1}},
        {"MgY", 3.21E-10, 3, 0, 18.0E9, 46.8E9,
         {126.3042E-3, -21.6452E-3, -29.3248E-3, 11.4053E-3, -1.2093E-3, -20.2756E-3, -10.9327E-3}, 30.0000E-3,
         83.0000E-3, 0, {}},
        {"Ti", 2.95E-10, 3, 0, 44.6E9, 118.67E9,
         {112.6205E-3, 0.7200E-3, -58.3781E-3, 28.1313E-3, -3.9212E-3, 25.2325E-3, -23.8545E-3}, 333.565E-3,
         622.5063E-3, 0,
         {427.0E-3, -165.9E-3, -221.9E-3, -26.29E-3, -13.42E-3, -53.09E-3, 119.1E-3, -14.25E-3, -17.77E-3}},
        {"MgAl", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {265.564638E-3, -96.111676E-3, 11.786691E-3, -7.169259E-3, 1.507026E-3, -161.871180E-3, 0.742991E-3},
         23.52941E-3, 83.83E-3, 0, {1}},
        {"MgCa", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {310.853100E-3, -114.443600E-3, 16.958600E-3, -11.221200E-3, 2.570400E-3, -191.432500E-3, 2.798000E-3},
         31.56863E-3, 100.39216E-3, 0, {1}},
        {"MgCe", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {241.210583E-3, -72.302347E-3, -13.746107E-3, 4.705126E-3, 0.523234E-3, 0.523234E-3, -7.398787E-3},
         1.17647E-3, 95.19608E-3, 0, {1}},
        {"MgGd", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {285.763113E-3, -95.960841E-3, 0.671163E-3, -1.588154E-3, 0.848439E-3, -159.569265E-3, -1.518973E-3},
         11.66667E-3, 102.64706E-3, 0, {1}},
        {"MgLi", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {281.067152E-3, -103.882445E-3, 16.615828E-3, -9.525039E-3, 1.569110E-3, -166.140969E-3, 4.570589E-3},
         49.01961E-3, 99.41176E-3, 0, {1}},
        {"MgSi", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {237.616625E-3, -85.750320E-3, 10.157547E-3, -5.707935E-3, 1.062990E-3, -148.577990E-3, -1.072479E-3},
         14.90196E-3, 66.86275E-3, 0, {1}},
        {"MgSn", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {265.383735E-3, -96.234971E-3, 12.703843E-3, -6.926203E-3, 1.011138E-3, -168.851904E-3, -0.639299E-3},
         12.25490E-3, 72.84314E-3, 0, {1}},
        {"MgZn", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {239.664844E-3, -83.074958E-3, 4.673015E-3, -3.589689E-3, 1.080065E-3, -132.943686E-3, -1.228621E-3},
         33.03922E-3, 85.09804E-3, 0, {1}},
        {"MgZr", 3.19E-10, 3, 0, 18.0E9, 46.8E9,
         {288.840762E-3, -104.774131E-3, 13.194615E-3, -8.018332E-3, 1.683542E-3, -168.501923E-3, 2.901528E-3},
         40.00000E-3, 100.78431E-3, 0, {1}},
        {"Ta", 3.3058E-10, 1, 0, 70.67E9, 189.24E9, {}, 0, 950.21E-3, 0, {}},
        {"W", 3.1652E-10, 1, 0, 160.6E9, 411.1E9, {}, 0, 1850.E-3, 0, {}},
        {"Cr", 2.8848E-10, 1, 0, 115.4E9, 279.3E9, {}, 0, 0, 0, {}},
        {"Fe", 2.8665E-10, 1, 0, 87.2796E9, 225.1814E9, {}, 0, 744.14E-3, 0, {}},
        {"Nb", 3.324E-10, 2, 2.859E-10, 39.64E9, 110.33E9, {}, 0, 676.775E-3, 676.775E-3, {}},
        {"Mo", 3.16E-10, 1, 0, 119.29E9, 310.68E9, {}, 0, 1.44339, 1.44339, {}},
        {"MoNbTi", 3.2373185E-10, 1, 0, 32.30E9, 90.70E9, {}, 0, 0.761408228239499, 0.761408228239499, {}},
        {"TaNbTi", 3.3014008E-10, 1, 0, 29.55E9, 82.71E9, {}, 0, 0.549789143600001, 0.549789143600001, {}},
        {"Nb_ML", 3.326E-10, 1, 0, 24.89E9, 71.52E9, {}, 0, 0.77774, 0.77774, {}},
        {"Mo_ML", 3.173E-10, 1, 0, 117.80E9, 305.67E9, {}, 0, 1.2321, 1.2321, {}},
        {"Ta_ML", 3.322E-10, 1, 0, 48.79E9, 134.60E9, {}, 0, 0.76308, 0.76308, {}},
    };
    for (const Row &r : rows) {
        if (name != r.name) continue;
        a = r.a;
        if (r.brule == 0) b = a / std::sqrt(2.0);
        else if (r.brule == 1) b = std::sqrt(3.0) * a / 2.0;
        else if (r.brule == 2) b = r.b;
        mu = r.mu;
        young = r.young;
        c0 = r.c[0], c1 = r.c[1], c2 = r.c[2], c3 = r.c[3], c4 = r.c[4], a1 = r.c[5], a3 = r.c[6];
        if (r.isf != 0) isf = r.isf;
        if (r.usf != 0) usf = r.usf;
        if (r.An != 0) An = r.An;
        // {1} marks the Mg pyramidal-II constants the Mg alloys reuse
        const double *py = (r.pyr[0] == 1 && r.pyr[1] == 0) ? MG_PYR : r.pyr;
        for (int q = 0; q < 9; ++q) pyr[q] = py[q];
        // conduction constants of the presets that have them (condsigma,
        // mfpel, alpha_resistivity, alpha_mfpel, Rgrain, beta_disloc)
        static const struct {
            const char *name;
            double v[6];
        } cond[] = {{"Ni", {1.43e7, 5.87e-9, 0.006, 0.006, 0.19, 10.e-25}},
                    {"Al", {3.77e7, 18.9e-9, 0.00429, 0.00429, 0.17, 1.5e-25}},
                    {"Au", {4.11e7, 37.7e-9, 0.0034, 0.0034, 0.17, 2.6e-25}},
                    {"Cu", {5.95e7, 39.9e-9, 0.00385, 0.00385, 0.24, 1.3e-25}},
                    {"Ag", {6.3e7, 53.3e-9, 0.0038, 0.0038, 0.24, 1.9e-25}},
                    {"Ta", {7.634e6, 3.21e-9, 0.0035, 0.0035, 0.1, 0}},
                    {"W", {1.855e7, 15.5e-9, 0.0042, 0.0042, 0.42, 7.5e-25}},
                    {"Cr", {7.937e6, 2.9e-9, 0.003, 0.003, 0.1, 0}},
                    {"Fe", {1.041e6, 40.0e-9, 0.005, 0.005, 0.1, 10.e-25}},
                    {"Nb", {6.334e6, 2.0e-9, 0.00318, 0.00318, 0.45, 0}}};
        for (const auto &cr : cond)
            if (name == cr.name) {
                condsigma = cr.v[0], mfpel = cr.v[1], alpha_resistivity = cr.v[2], alpha_mfpel = cr.v[3];
                Rgrain = cr.v[4];
                if (cr.v[5] != 0) beta_disloc = cr.v[5];
            }
        // USFE(angle, tau) fits of core flag 6 (material.cpp:713-717, 766-770)
        if (name == "Ta") a_slope = 51.59816407, a_b = -2.27173784, b_b = 0.9979452, c_slope = -4.16536, c_b = 1133.961364;
        if (name == "Fe")
            a_slope = 15.524609942496577, a_b = 0.004388435835439675, b_b = 0.9934680885709689,
            c_slope = -0.9065590784654822, c_b = 944.0633238797747;
        return true;
    }
    return false;
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
        t.default_init = "mode1";
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
        t.default_init = "3D3ShcpBasal";
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
    if (app == "1L1S100B" || app == "cond100B") {
        // App1L1S100B (app_1L1S_100B.cpp:29-49, 215-229) and Appcond100B
        // (app_cond_100B.cpp:29-49, 90-104): one screw system, b = x on the
        // plane n = z, no rotation block; default setup StraightNonOrtho
        t.slip_systems = 1;
        t.num_planes = 1;
        t.CD = 0.5;
        t.default_core = 1;
        t.default_init = "StraightNonOrtho";
        t.rotate = false;
        t.xn = {0.0, 0.0, 1.0};
        t.xb = {1.0, 0.0, 0.0};
        return true;
    }
    if (app == "2d-bcc-1lOrtho") {
        // App2Dbcc1LORTHO (app_2d_bcc_ortho_1loop.cpp:29-49, 210-256):
        // n = (1, -1, 0)/sqrt2, b = (1, 1, 1)/2, rotated by four/normfour;
        // default setup 1LOrthoBCC
        t.slip_systems = 1;
        t.num_planes = 1;
        t.CD = 0.5;
        t.default_core = 1;
        t.default_init = "1LOrthoBCC";
        t.xn = {1.0 / std::sqrt(2), -1.0 / std::sqrt(2), 0.0};
        t.xb = {0.5, 0.5, 0.5};
        return true;
    }
    {
        // The single-table apps of the reference: slip systems, mobility,
        // whether set_slip rotates by four/normfour, and the default setup
        // (setup_flag of the constructor; "" = that initial configuration is
        // not built here).  Vectors as set_slip writes them.
        const double r2 = 1.0 / std::sqrt(2.0), r3 = 1.0 / std::sqrt(3.0);
        struct AppRow {
            const char *name;
            int S;
            double CD;
            bool rotate;
            const char *init, *missing; // built default setup / name of an unbuilt one
            double xn[9], xb[9];
        };
        const AppRow rows[] = {
            // app_1L1S111B_bcc.cpp:29-49, 212-228 (rotation block commented out)
            {"1L1S111Bbcc", 1, 0.5, false, "1LOrthoBCC", "", {0, 0, 1}, {1, 0, 0}},
            // app_1L1S_110B_fcc.cpp, _fcc_nonortho.cpp, _inclined.cpp, app_1L1S_111B.cpp
            {"1L1S110Bfcc", 1, 0.5, false, "1LOrthoFCC", "", {r3, r3, r3}, {-r2, r2, 0}},
            {"1L1S110Bfccnonortho", 1, 0.5, false, "1LNonOrthoFCC", "", {r3, r3, r3}, {r2, -r2, 0}},
            {"1L1S110Binclined", 1, 0.5, false, "1LInclinedFCC", "", {r3, r3, r3}, {r2, -r2, 0}},
            {"1L1S111B", 1, 0.5, false, "1LOrthoFCC", "", {r3, r3, r3}, {-0.5, 0.5, 0}},
            // app_1SBCC.cpp, app_3SBCC.cpp, app_bcc_edge.cpp, app_bcc_edge_planar.cpp
            {"1SBCC", 1, 0.25, false, "3SBCC", "", {0.7071, 0.7071, 0}, {-0.57735, 0.57735, 0.57735}},
            {"3SBCC", 3, 0.25, false, "3SBCC", "", {0.7071, 0.7071, 0, 0, 0.7071, -0.7071, 0.7071, 0, 0.7071},
             {-0.57735, 0.57735, 0.57735, -0.57735, 0.57735, 0.57735, -0.57735, 0.57735, 0.57735}},
            {"BCCEdge", 3, 0.25, false, "edge", "", {0.7071, 0.7071, 0, 0, 0.7071, -0.7071, 0.7071, 0, 0.7071},
             {-0.57735, 0.57735, 0.57735, -0.57735, 0.57735, 0.57735, -0.57735, 0.57735, 0.57735}},
            {"BCCEdgePlanar", 1, 0.25, false, "edge", "", {0.7071, 0.7071, 0}, {-0.57735, 0.57735, 0.57735}},
            // app_2L2S_110B.cpp (no rotation), app_2d.cpp (rotated)
            {"2L2S110B", 2, 0.5, false, "mode1", "", {r3, r3, r3, r3, r3, -r3}, {r2, -r2, 0, r2, -r2, 0}},
            {"2d", 1, 0.5, true, "1LOrthoFCC", "", {r3, r3, r3}, {0.5, -0.5, 0}},
            // app_3d_2slip*.cpp (rotated)
            {"3d-2slip", 1, 0.5, true, "1LOrthoBCC", "", {r3, r3, r3}, {-0.5, 0.5, 0}},
            {"3d2s2lOrtho", 2, 0.5, true, "1LOrthoBCC", "", {r3, r3, r3, r3, r3, -r3}, {0.5, -0.5, 0, 0.5, -0.5, 0}},
            {"3d2sStraightNonOrtho", 2, 0.5, true, "StraightNonOrtho", "", {-r3, r3, -r3, r3, -r3, -r3},
             {0.5, 0.5, 0, 0.5, 0.5, 0}},
            {"3d2sStraightOrtho", 2, 0.5, true, "StraightOrtho", "", {r3, r3, r3, r3, r3, -r3},
             {0.5, -0.5, 0, 0.5, -0.5, 0}},
        };
        for (const AppRow &r : rows) {
            if (app != r.name) continue;
            t.slip_systems = r.S;
            t.num_planes = r.S;
            t.CD = r.CD;
            t.default_core = 1;
            t.rotate = r.rotate;
            t.default_init = r.init;
            t.missing_init = r.missing;
            t.xn.assign(r.xn, r.xn + 3 * r.S);
            t.xb.assign(r.xb, r.xb + 3 * r.S);
            return true;
        }
        if (app == "3d3sHCPnotch") {
            // App3d3sHCPnotch (app_3d3s_hcp_notch.cpp:29-50, set_slip): the
            // 3d-hcp-basal vectors with oflag = 0, rotated; default setup
            // 3D3ShcpNotch
            t.slip_systems = 3;
            t.num_planes = 3;
            t.CD = 0.5;
            t.default_core = 1;
            t.default_init = "3D3ShcpNotch";
            const double ang[3] = {0.0, M_PI / 3.0, -M_PI / 3.0};
            for (int a = 0; a < 3; ++a) {
                t.xn.insert(t.xn.end(), {0.0, 0.0, 1.0});
                t.xb.insert(t.xb.end(), {1.0 * std::cos(ang[a]), 1.0 * std::sin(ang[a]), 0.0});
            }
            return true;
        }
    }
    if (app == "frank_read" || app == "FrankRead") {
        // AppFrankRead (app_frank_read.cpp:29-50, 96-120); its init_app always
        // sets up initial_sxtal_frank_read (:77-85, 127-130)
        t.default_init = "frank_read";
        t.force_init = app == "FrankRead";
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

void resolss_non_schmid(const SlipTable &t, const double sigma_rot[3][3], double dslip, double angle, double w1,
                        double w2, double w3, double *tau)
{
    // FFTW_Slab::resolSS_non_Schmid (fft_fftw_slab.cpp:774-845; R. Groger et
    // al., Acta Mater. 56 (2008)): per system the maximum-resolved-shear plane
    // n (angle_to_110 + 120 deg per system about y), the plane n' at +60 deg,
    // and tau = sigma_rot : [b n + w1 b n' + w2 (n x b) n + w3 (n' x b) n'] / dslip.
    double theta = 0.0;
    for (int is = 0; is < t.slip_systems; ++is, theta += 120.0) {
        const double n[3] = {-std::sin((theta + angle) * M_PI / 180.), 0.0, std::cos((theta + angle) * M_PI / 180.)};
        const double np[3] = {-std::sin((theta + angle + 60) * M_PI / 180.), 0.0,
                              std::cos((theta + angle + 60) * M_PI / 180.)};
        const double *b = &t.xb[3 * is];
        const double nxb[3] = {n[1] * b[2] - n[2] * b[1], -n[0] * b[2] + n[2] * b[0], n[0] * b[1] - n[1] * b[0]};
        const double npxb[3] = {np[1] * b[2] - np[2] * b[1], -np[0] * b[2] + np[2] * b[0], np[0] * b[1] - np[1] * b[0]};
        double v = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double S = n[j] * b[i] + w1 * np[j] * b[i] + w2 * n[j] * nxb[i] + w3 * np[j] * npxb[i];
                v = v + (sigma_rot[i][j] * S) / dslip;
            }
        tau[is] = v;
    }
}

double usfe_angle_tau(const Material &m, const double sigma_rot[3][3], double angle)
{
    // core_energy_USFE_angle_tau (fft_fftw_slab.cpp:1500-1545): the unstable
    // stacking-fault energy as a function of the MRSSP angle and the resolved
    // shear sigma_rot[2][1] (mJ/m^2 fit), returned as An' = gsfe/(mu dslip b)
    const double s21 = sigma_rot[2][1] * m.mu_s * 1.0E-9;
    double g = (m.a_slope * s21 + m.a_b) * std::sin(m.b_b * angle * M_PI / 180.) + (m.c_slope * s21 + m.c_b);
    g *= 1.0E-3;
    return g / (m.mu_s * m.dslip * m.b);
}

void project_core_energy(int nx, int ny, int nz, int ns, const double *xb, const double *xi, double *delta,
                         double *ddelta)
{
    const long long N = (long long)nx * ny * nz;
    for (long long r = 0; r < N; ++r) delta[r] = ddelta[r] = 0.0;
    for (long long r = 0; r < N; ++r) {
        if (ns == 1 || ns == 2) {
            delta[r] = xi[2 * r];
        } else {
            double v = 0.0;
            for (int c = 0; c < 3; ++c) {
                double p = 0.0;
                for (int s = 0; s < 3; ++s) p += xi[2 * (r + s * N)] * xb[3 * s + c];
                v += p * xb[3 * 1 + c];
            }
            delta[r] = v;
        }
    }
    const int j = ny / 2, k = nz / 2;
    for (int i = 0; i < nx; ++i) {
        const long long q = (long long)i * ny * nz + (long long)j * nz + k;
        ddelta[q] = i + 1 == nx ? 0.0 : (delta[q + (long long)ny * nz] - delta[q]) / 1.0;
    }
}

std::vector<double> core_coefficients(const Material &m, int flag)
{
    const double mdb = m.mu_s * m.dslip * m.b;
    if (flag == 1) return {m.An / mdb};
    if (flag == 5) return {m.usf / mdb}; // core_energy_bcc_perfect (:1090-1094)
    if (flag == 2) {
        // core_energy_extended (fft_fftw_slab.cpp:1281-1287, 1318): c/mu, density/(dslip b)
        return {m.c0 / mdb, m.c1 / mdb, m.c2 / mdb, m.c3 / mdb, m.c4 / mdb, m.a1 / mdb, m.a3 / mdb};
    }
    if (flag == 4 || flag == 9) {
        // core_energy_pyrII (:1468-1476): aa0..aa4, bb1..bb4 over (mu dslip b)
        std::vector<double> c(9);
        for (int q = 0; q < 9; ++q) c[q] = m.pyr[q] / mdb;
        return c;
    }
    if (flag == 3) {
        const double isf = m.isf / mdb;
        const double Cn = (m.usf - (isf / 2.0)) / mdb;
        return {isf, Cn};
    }
    return {};
}

bool cond_layers(const Material &A, const Material &B, int ny, double temperature, std::vector<double> &sigma,
                 std::vector<double> &dsigma, std::string *msg)
{
    // backwardcond, fft_fftw_slab.cpp:2083-2159: alternating layers A | B
    // along y, starting with A; integer division N2/Nlayers as written
    const double Vf = B.dlayer / (A.dlayer + B.dlayer);
    const int Nl = A.Nlayers;
    const double prob = 1.0 - A.pinterface;
    const double dT = temperature - 300;
    if (Nl < 1 || ny < 2) {
        if (msg) *msg = "conductivity needs Nlayers >= 1 and ny >= 2";
        return false;
    }
    const int dB = (int)std::round((double)(ny / Nl) * Vf);
    const int dA = (int)std::round((double)(ny / Nl) * (1. - Vf));
    if (dA + dB < 1) {
        if (msg) *msg = "layer thicknesses round to zero rows";
        return false;
    }
    std::vector<double> bulk(ny);
    sigma.assign(ny, 0.0);
    for (int j = 0; j < ny; ++j) {
        const Material &m = (j % (dA + dB) >= dA) ? B : A;
        bulk[j] = m.condsigma / (1. + m.alpha_resistivity * dT);
        const double PSprob = 2 * prob / m.dlayer;
        const double lam = m.mfpel / (1. + m.alpha_mfpel * dT);
        const double ag = (lam * m.Ngrains / m.dlayer) * m.Rgrain / (1 - m.Rgrain);
        const double rhod = m.beta_disloc * m.rho_disloc;
        double cs;
        if (PSprob > 0.) {
            const double rhoi = (1 / bulk[j]) * ((3. / 16.) * lam * PSprob);
            if (m.Ngrains > 1) {
                const double rhog =
                    (1 / bulk[j]) * (1. / (1. - 3. * ag / 2. + 3. * ag * ag - 3. * std::pow(ag, 3.) * std::log(1. + 1. / ag)) - 1.);
                cs = 1 / ((1 / bulk[j]) + (2 * rhoi + (m.Ngrains - 1) * rhog) / (m.Ngrains + 1) + rhod);
            } else {
                cs = 1 / ((1 / bulk[j]) + rhoi + rhod);
            }
        } else {
            cs = bulk[j];
        }
        sigma[j] = cs;
    }
    dsigma.assign(ny, 0.0);
    for (int j = 0; j < ny; ++j) {
        if (j == 0) dsigma[j] = bulk[j + 1] - bulk[j];
        else if (j == ny - 1) dsigma[j] = bulk[j] - bulk[j - 1];
        else dsigma[j] = (bulk[j + 1] - bulk[j - 1]) / 2.;
    }
    return true;
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

void init_sxtal_straight_nonortho(const Grid &g, double *xi)
{
    // initial_sxtal_StraightNonOrtho (fft_fftw_slab.cpp:2804-2855): system 0,
    // the plane k = nz/2 (|k - nz/2| < 0.1), cells with 2i - j <= -nx or
    // 0 <= 2i - j <= nx
    const long long N = g.N();
    const double eps = 1e-1;
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const double dk = k - g.nz / 2.0;
                    const int d = 2 * i - j;
                    const bool on = dk < eps && dk > -eps && (d <= -g.nx || (d >= 0 && d <= g.nx)) && is == 0;
                    xi[2 * idx] = on ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                }
}

bool init_sxtal_more(const std::string &style, const Grid &g, const Geometry &geo, double sa, double xprd, double *xi,
                     double *xo)
{
    // The remaining hard-coded configurations of FFTW_Slab (fft_fftw_slab.cpp):
    // mode1 :2245-2333, mode2 :2336-2410, 1LOrthoFCC :2413-2471,
    // 1LNonOrthoFCC :2474-2537, 1LInclinedFCC :2540-2597, 2LOrtho :2665-2742,
    // 3D3ShcpBasal :2864-2946, 3D3ShcpNotch :2949-3029 (also the mask xo).
    // Cell (i, j, k) of system is; the plane tests |.| < eps are the
    // reference's; sclprim maps grid offsets to lengths (mode1, 1LNonOrthoFCC).
    const long long N = g.N();
    const int N1 = g.nx, N2 = g.ny, N3 = g.nz;
    auto len = [&](double a, double b, double c) {
        double r2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double v = geo.sclprim[q][0] * a + geo.sclprim[q][1] * b + geo.sclprim[q][2] * c;
            r2 += v * v;
        }
        return std::sqrt(r2);
    };
    int which;
    if (style == "mode1") which = 0;
    else if (style == "mode2") which = 1;
    else if (style == "1LOrthoFCC") which = 2;
    else if (style == "1LNonOrthoFCC") which = 3;
    else if (style == "1LInclinedFCC") which = 4;
    else if (style == "2LOrtho") which = 5;
    else if (style == "3D3ShcpBasal") which = 6;
    else if (style == "3D3ShcpNotch") which = 7;
    else return false;
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < N1; ++i)
            for (int j = 0; j < N2; ++j)
                for (int k = 0; k < N3; ++k) {
                    const long long idx = k + (long long)j * N3 + (long long)i * N2 * N3 + (long long)is * N;
                    bool on = false, mask = false;
                    const double dk2 = k - N3 / 2.0;
                    switch (which) {
                    case 0: {
                        const double eps = 1e-1, dj2 = j - N2 / 2.0;
                        if (dk2 < eps && dk2 > -eps && is == 0)
                            on = len(i - N1 / 2.0, j - 5.0 * N2 / 8.0, k - N3 / 2.0) <= N2 / 8.0 - 2.0;
                        else if (dj2 < eps && dj2 > -eps && is == 1)
                            on = len(i - N1 / 2.0 + N1 / 8.0, j - N2 / 2.0, k - 5.0 * N3 / 8.0) <= N3 / 8.0 - 2.0;
                        break;
                    }
                    case 1: {
                        const double eps = 1e-1, p1 = i + j + k - N1, p2 = i + j - k;
                        if (p1 < eps && p1 > -eps && is == 0)
                            on = sa * std::sqrt((i - 28.0) * (i - 28.0) + (j - 28.0) * (j - 28.0) + (k - 40.0) * (k - 40.0)) <=
                                 3.0 * std::sqrt(2.0);
                        else if (p2 < eps && p2 > -eps && is == 1)
                            on = sa * std::sqrt((i - 28.0) * (i - 28.0) + (j - 28.0) * (j - 28.0) + (k - 56.0) * (k - 56.0)) <=
                                 3.0 * std::sqrt(2.0);
                        break;
                    }
                    case 2: {
                        const double eps = 1e-1;
                        if (dk2 < eps && dk2 > -eps && is == 0)
                            on = std::sqrt((i - N1 / 2.0) * (i - N1 / 2.0) + (j - N2 / 2.0) * (j - N2 / 2.0)) <= xprd / 10.0;
                        break;
                    }
                    case 3: {
                        const double eps = 1e-1;
                        if (dk2 < eps && dk2 > -eps && is == 0)
                            on = len(i - N1 / 2.0, j - N2 / 2.0, k - N3 / 2.0) <= 10.0;
                        break;
                    }
                    case 4: {
                        const double eps = 1e-1, p1 = i + j + k - N1;
                        if (p1 < eps && p1 > -eps && is == 0)
                            on = sa * std::sqrt((i - N1 / 3.0) * (i - N1 / 3.0) + (j - N2 / 3.0) * (j - N2 / 3.0) +
                                                (k - N3 / 3.0) * (k - N3 / 3.0)) <= 5.0 * std::sqrt(2.0);
                        break;
                    }
                    case 5: {
                        const double eps = 1e-1, t = 2.8284 * (j - N2 / 2.0) + (k - N3 / 2.0), dk4 = k - N3 / 4.0;
                        if (t < 20.0 * eps && t > -20.0 * eps && is == 1)
                            on = std::sqrt((i - N1 / 2.0) * (i - N1 / 2.0) + (j - N2 / 2.0) * (j - N2 / 2.0) +
                                           (k - N3 / 2.0) * (k - N3 / 2.0)) <= xprd / 20.0;
                        else if (dk4 < eps && dk4 > -eps && is == 0)
                            on = std::sqrt((i - N1 / 2.0) * (i - N1 / 2.0) + (j - N2 / 2.0) * (j - N2 / 2.0) +
                                           (k - N3 / 4.0) * (k - N3 / 4.0)) <= xprd / 20.0;
                        break;
                    }
                    case 6: {
                        const double eps = 1e-2;
                        on = dk2 < eps && dk2 > -eps && i > N1 / 4.0 && i < 3.0 * N1 / 4.0 && is == 0;
                        break;
                    }
                    case 7: {
                        const double eps = 1e-2;
                        mask = i < 5 || j < 5 || k < 5;
                        if (dk2 < eps && dk2 > -eps) {
                            const double ir = (i - 5.0) * (i - 5.0) + (double)((j - N2 / 2) * (j - N2 / 2));
                            if (ir <= 25 && (i == 5 || i == 6)) {
                                mask = true;
                                on = is == 0;
                            }
                        }
                        break;
                    }
                    }
                    xi[2 * idx] = on ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                    if (xo) xo[idx] = mask ? 1.0 : 0.0;
                }
    return true;
}

void init_sxtal_straight_ortho(const Grid &g, double *xi)
{
    // initial_sxtal_StraightOrtho (fft_fftw_slab.cpp:2745-2800): system 0,
    // the plane |k - nz/2| < 0.2, rows ny/4 < j < 3 ny/4
    const long long N = g.N();
    const double eps = 2e-1;
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const double dk = k - g.nz / 2.0;
                    const bool on = dk < eps && dk > -eps && j > g.ny / 4.0 && j < 3.0 * g.ny / 4.0 && is == 0;
                    xi[2 * idx] = on ? 1.0 : 0.0;
                    xi[2 * idx + 1] = 0.0;
                }
}

void init_sxtal_1lorthobcc(const Grid &g, double sa, double sb, double *xi)
{
    // initial_sxtal_1LOrthoBCC (fft_fftw_slab.cpp:2600-2660): system 0, the
    // plane k = nz/2, a disc of radius 10 sa about (nx/2, ny/2) in the
    // scaled coordinates (sa (i - nx/2), sb (j - ny/2))
    const long long N = g.N();
    const double eps = 1e-1;
    for (int is = 0; is < g.S; ++is)
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)is * N;
                    const double dk = k - (g.nz / 2.0);
                    bool on = false;
                    if (dk < eps && dk > -eps && is == 0) {
                        const double ir = std::sqrt(sa * (i - g.nx / 2.0) * sa * (i - g.nx / 2.0) +
                                                    sb * (j - g.ny / 2.0) * sb * (j - g.ny / 2.0));
                        on = ir <= 10 * sa;
                    }
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
