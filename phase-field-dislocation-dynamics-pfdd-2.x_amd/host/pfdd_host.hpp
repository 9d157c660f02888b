// pfdd_host.hpp -- host-side (C++) mirror of the PFDD-2.x setup that feeds the
// GL solver: material constants, primitive/reciprocal lattice, slip tables of
// the app styles, stress rotation and resolved shear, RANMAR and the initial
// order-parameter fields.  Everything here runs once per run or once per
// stress state; the per-step work is on the GPU (include/pfdd_gl.h).
#pragma once

#include <string>
#include <vector>

namespace pfdd_host {

// Material::Material + set_constants (material.cpp:32-167, presets :176-835).
struct Material {
    double a = 0, b = 0, mu = 0, young = 0;
    double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, a1 = 0, a3 = 0;
    double isf = 0, usf = 0, An = 0, Cn = 0;
    double pyr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}; // aa0..aa4, bb1..bb4 (pyramidal II)
    double a_slope = 0, a_b = 0, b_b = 0, c_slope = 0, c_b = 0; // USFE(angle, tau) fit (flag 6)
    double dslip = 0.8165;
    // electrical conduction (material.h:44-56, presets of material.cpp), read by
    // the conductivity sub-problem of solve_style GLCond
    double condsigma = 0, alpha_resistivity = 0, Rgrain = 0, mfpel = 1e-9, alpha_mfpel = 0;
    double dlayer = 100e-9, pinterface = 0.5, epsilonzero = 8.8541878128e-12;
    double beta_disloc = 0, rho_disloc = 0;
    int Nlayers = 1, Ngrains = 1;
    double El[3] = {0, 0, 0};
    // derived (material.cpp:156-167)
    double C44 = 0, C12 = 0, C11 = 0, nu = 0, ll = 0, S11 = 0, S12 = 0, S44 = 0;
    // Voigt re-derivation stored by Bmatrix (fft_fftw_slab.cpp:487-496)
    double mu_s = 0, ll_s = 0, young_s = 0, xnu_s = 0;
    std::vector<double> A_mpea; // per-cell USFE of core flag 7 (J/m^2), [a][i][j][k]
    int mpea_dims[4] = {0, 0, 0, 0}; // n1 n2 n3 ns of the MPEA file header

    bool preset(const std::string &name); // the named presets of material.cpp
    void derive();
};

// Primitive vectors (fft.cpp:96-169) and reciprocal matrix (app.cpp:383-431).
struct Geometry {
    double prim[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double invprim[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double norvol = 1.0;
    double sclprim[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double four[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    double normfour[3] = {1, 1, 1};
    double vol = 1.0;
    bool primitive = false;
    void set_primitive(const double x[3], const double y[3], const double z[3]);
    void set_scale(double sa, double sb, double sc);
};

// Slip systems of an app style, before rotation into the grid frame.
struct SlipTable {
    std::string app;
    int slip_systems = 0;
    int num_planes = 0;
    double CD = 0.5;       // dislocation mobility (app ctor)
    double timestep = 1.0; // app->timestep
    int default_core = 1;
    std::string default_init; // the app's default setup_flag as an initial style ("" = none built)
    bool force_init = false;  // the app's init_app ignores the deck's setup (AppFrankRead)
    std::string missing_init; // the app's default setup when it is not built here
    bool rotate = true;         // set_slip applies the four/normfour rotation
    std::vector<double> xn, xb; // S*3 each
};

// Known app styles: every AppStyle of the reference (fcc, 3d-hcp-basal,
// FrankRead (alias frank_read), 1L1S100B, cond100B, 2d-bcc-1lOrtho,
// 1L1S111Bbcc, 1L1S110Bfcc, 1L1S110Bfccnonortho, 1L1S110Binclined, 1L1S111B,
// 1SBCC, 3SBCC, BCCEdge, BCCEdgePlanar, 2L2S110B, 2d, 3d-2slip, 3d2s2lOrtho,
// 3d2sStraightNonOrtho, 3d2sStraightOrtho, 3d3sHCPnotch) and basal3 (three
// unit b at 0/120/240 deg on the plane n = z).  Returns false otherwise.
bool slip_table(const std::string &app, SlipTable &t, double oflag = 0.0);

// AppFCC::set_slip rotation block (app_fcc.cpp:288-313).
void rotate_slip(SlipTable &t, const Geometry &g);
// FFTW_Slab::rotate_stress (fft_fftw_slab.cpp:242-274).
void rotate_stress(const Geometry &g, const double sigma[3][3], double sigma_rot[3][3]);
// FFTW_Slab::resolSS_Schmid (fft_fftw_slab.cpp:737-769).
void resolss(const SlipTable &t, const double sigma_rot[3][3], double dslip, double *tau);
// FFTW_Slab::resolSS_non_Schmid (fft_fftw_slab.cpp:774-845), the non_Schmid
// command's angle_to_110 and weights w1..w3 (app.cpp:446-460).
void resolss_non_schmid(const SlipTable &t, const double sigma_rot[3][3], double dslip, double angle, double w1,
                        double w2, double w3, double *tau);
// An' of core_energy_USFE_angle_tau (fft_fftw_slab.cpp:1500-1545).
double usfe_angle_tau(const Material &m, const double sigma_rot[3][3], double angle);

// project_core_energy (fft_fftw_slab.cpp:1550-1624) for plane 0, the one the
// apps' compute_delta / compute_ddelta read: delta = Re xi_0 for one or two
// systems, else the projection of the plane's three systems onto b_1;
// ddelta = forward difference along x on the line j = ny/2, k = nz/2 (dx = 1),
// zero at the last cell and off the line.  delta, ddelta: N values, zeroed
// here first; xi in the fft->xi[0] layout; xb: ns*3.
void project_core_energy(int nx, int ny, int nz, int ns, const double *xb, const double *xi, double *delta,
                         double *ddelta);

// Core-energy coefficients normalised as core_energy_* do (fft_fftw_slab.cpp:
// 1165-1178, 1396-1406): flag 1 -> {An'}, flag 3 -> {isf', Cn'} with Cn' =
// (usf - isf'/2)/(mu dslip b), the reference's own expression.
std::vector<double> core_coefficients(const Material &m, int flag);

// Per-row conductivity of the two-material layer stack of backwardcond
// (fft_fftw_slab.cpp:2083-2159, rows j along y): A = the first material
// command (fft->material), B = the second; sigma[j] = condsigma of row j
// (bulk / (1 + alpha dT) with interface, grain-boundary and dislocation
// resistivities added), dsigma[j] = the difference of the bulk values across
// row j (:2171-2182).  Returns false (and leaves msg) on a stack with no rows.
bool cond_layers(const Material &A, const Material &B, int ny, double temperature, std::vector<double> &sigma,
                 std::vector<double> &dsigma, std::string *msg = nullptr);

// RanMars (random_mars.cpp:41-101).
class RanMars {
  public:
    explicit RanMars(int seed);
    double uniform();

  private:
    double u[98];
    double c, cd, cm;
    int i97, j97;
};

// Initial fields, reference layout xi[2*(k + j*nz + i*ny*nz + a*N) + {re,im}].
struct Grid {
    int nx, ny, nz, S;
    long long N() const { return (long long)nx * ny * nz; }
};
// Seeded loops: Re xi_a = 1 on the cells of a Miller-plane disc
// {h i + k j + l k = m0, |r - c| <= R} plus noise 2*amp*(u - 1/2) on every
// cell (RANMAR, order a, i, j, k), then the loop draws from the same stream.
void init_random_loops(const Grid &g, const SlipTable &t, int nloops, double rmin, double rmax,
                       double noise, int seed, double *xi, std::vector<double> *loops_out = nullptr);
// initial_sxtal_loop (fft_fftw_slab.cpp:3146-3197).
void init_sxtal_loop(const Grid &g, double *xi);
// initial_sxtal_frank_read (fft_fftw_slab.cpp:3032-3093): xi and the mask xo.
void init_sxtal_frank_read(const Grid &g, double *xi, double *xo);
// initial_sxtal_edge (fft_fftw_slab.cpp:3201-3247).
void init_sxtal_edge(const Grid &g, double *xi);
// initial_sxtal_StraightNonOrtho (fft_fftw_slab.cpp:2804-2855).
void init_sxtal_straight_nonortho(const Grid &g, double *xi);
// mode1, mode2, 1LOrthoFCC, 1LNonOrthoFCC, 1LInclinedFCC, 2LOrtho, 3D3ShcpBasal,
// 3D3ShcpNotch (fft_fftw_slab.cpp:2245-2742, 2864-3029); xo (may be NULL) gets
// the notch's mask.  sa = scale a, xprd = box length along x.  False for
// other names.
bool init_sxtal_more(const std::string &style, const Grid &g, const Geometry &geo, double sa, double xprd, double *xi,
                     double *xo);
// initial_sxtal_StraightOrtho (fft_fftw_slab.cpp:2745-2800).
void init_sxtal_straight_ortho(const Grid &g, double *xi);
// initial_sxtal_1LOrthoBCC (fft_fftw_slab.cpp:2600-2660); sa, sb = scale command.
void init_sxtal_1lorthobcc(const Grid &g, double sa, double sb, double *xi);
// initial_sxtal_3SBCC (fft_fftw_slab.cpp:3095-3144).
void init_sxtal_3sbcc(const Grid &g, double *xi);

} // namespace pfdd_host
