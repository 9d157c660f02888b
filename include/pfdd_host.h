/*
 * pfdd_host.h -- C ABI of the host-side setup helpers that feed the GL solver
 * (include/pfdd_gl.h): material constants, slip tables of the app styles,
 * stress rotation and resolved shear, RANMAR and initial fields.  Host code
 * only (no GPU).  Each function names the reference code it restates
 * (lanl/Phase-Field-Dislocation-Dynamics-PFDD-2.x, src/ directory).
 * Functions return 0 on success, nonzero on an invalid argument.
 */
#ifndef PFDD_HOST_H
#define PFDD_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfdd_host_material {
    double a, b, mu, young, isf, usf, An, Cn; /* inputs (material command)          */
    double dslip;                             /* 0.8165, material.cpp:166           */
    double C44, C12, C11, nu, S11, S12, S44;  /* derived, material.cpp:156-164      */
    double mu_s, xnu_s;                       /* Bmatrix Voigt mu, nu (:487-496)    */
} pfdd_host_material;

/* Material::set_constants preset ("Cu", "Ni", "Pd"; NULL or "" = none),
 * material.cpp:176-276.  Fields not set by the preset are zero. */
int pfdd_host_material_preset(const char *preset, pfdd_host_material *m);
/* Material::Material tail, material.cpp:156-167, plus the Voigt moduli. */
int pfdd_host_material_derive(pfdd_host_material *m);

/* Slip table of an app style ("fcc", "3d-hcp-basal", "basal3", "frank_read"),
 * rotated into the grid frame as the app's set_slip does (app_fcc.cpp:
 * 211-313).  prim9 = primitive x,y,z vectors (NULL = identity; fft.cpp:96-
 * 133), scale3 = the scale command (NULL = 1 1 1; app.cpp:383-431).
 * Outputs: *ns, xn/xb (12*3 capacity), four9 (row-major), *cd, *timestep. */
int pfdd_host_slip(const char *app, const double *prim9, const double *scale3, int *ns, double *xn,
                   double *xb, double *four9, double *cd, double *timestep);

/* rotate_stress + resolSS_Schmid (fft_fftw_slab.cpp:242-274, 737-769):
 * sigma9 in units of mu (sigma command), tau[ns], sigma_rot9. */
int pfdd_host_resolss(const char *app, const double *prim9, const double *scale3, const double *sigma9,
                      double dslip, double *tau, double *sigma_rot9);

/* rotate_stress + resolSS_non_Schmid (fft_fftw_slab.cpp:242-274, 774-845),
 * the non_Schmid command's angle_to_110 (deg) and weights w1, w2, w3. */
int pfdd_host_resolss_non_schmid(const char *app, const double *prim9, const double *scale3,
                                 const double *sigma9, double dslip, double angle, double w1, double w2, double w3,
                                 double *tau);

/* core_energy_* coefficient normalisation (fft_fftw_slab.cpp:1165-1178,
 * 1396-1406): flag 1 -> coef[0] = An'; flag 3 -> coef[0..1] = isf', Cn'.
 * Returns the number of coefficients in *ncoef. */
int pfdd_host_core_coef(const pfdd_host_material *m, int flag, double *coef, int *ncoef);

/* RanMars stream (random_mars.cpp:41-101): n uniforms. */
int pfdd_host_ranmar_fill(int seed, long long n, double *out);

/* Seeded initial field for an app's slip table: uniform noise 2*amp*(u-1/2)
 * on every system/cell (RANMAR, order a,i,j,k), then nloops Miller-plane
 * discs of radius in [rmin, rmax] (plane, system, centre, radius drawn from
 * the same stream).  xi: 2*ns*N doubles in the fft->xi[0] layout.  loops
 * (5*nloops, may be NULL) receives system, centre i,j,k, radius. */
int pfdd_host_init_random_loops(int nx, int ny, int nz, const char *app, int nloops, double rmin,
                                double rmax, double noise, int seed, double *xi, double *loops);

/* The reference's hard-coded initial conditions: style "loop"
 * (initial_sxtal_loop :3146-3197), "frank_read" (:3032-3093, also fills the
 * obstacle mask xo if non-NULL), "edge" (:3201-3247), "3SBCC" (:3095-3144),
 * "StraightNonOrtho" (:2804-2855), "StraightOrtho" (:2745-2800), "1LOrthoBCC"
 * (:2600-2660), "mode1", "mode2", "1LOrthoFCC", "1LNonOrthoFCC",
 * "1LInclinedFCC", "2LOrtho", "3D3ShcpBasal", "3D3ShcpNotch" (:2245-2742,
 * 2864-3029; the notch also fills xo) -- on the cubic grid, scale 1 1 1. */
int pfdd_host_init_sxtal(const char *style, int nx, int ny, int nz, int ns, double *xi, double *xo);

/* Conduction constants of a material (material.h:44-56 defaults, then the
 * preset's values from material.cpp; NULL or "" = defaults only) as par[11]
 * = condsigma, mfpel, alpha_resistivity, alpha_mfpel, Rgrain, beta_disloc,
 * rho_disloc, dlayer, Nlayers, Ngrains, pinterface. */
int pfdd_host_cond_preset(const char *preset, double *par);
/* Per-row conductivities of backwardcond's two-material layer stack
 * (fft_fftw_slab.cpp:2083-2159, 2171-2182), materials A and B as par[11]
 * above: sigma[ny] (condsigma of row j), dsigma[ny] (the bulk difference
 * across row j).  These are pfdd_gl_cond_desc's sigma / dsigma. */
int pfdd_host_cond_layers(const double *parA, const double *parB, int ny, double temperature, double *sigma,
                          double *dsigma);

/* The same formulas on the caller's own arrays, for code that already holds
 * the reference's geometry and slip tables (the fft_style hip plugin):
 * rotate_stress (fft_fftw_slab.cpp:242-274) from prim / invprim (row-major)
 * and norvol (fft.cpp:101-169); resolSS_Schmid (:737-769) and
 * resolSS_non_Schmid (:774-845) from the rotated slip vectors xn, xb (ns*3);
 * core_energy_USFE_angle_tau's An' (:1500-1545) with m->mu_s, dslip, b and the
 * fit par5 = a_slope, a_b, b_b, c_slope, c_b; project_core_energy
 * (:1550-1624) of plane 0 from xi (fft->xi[0] layout) into delta, ddelta (N). */
int pfdd_host_rotate_stress_raw(const double *prim9, const double *invprim9, double norvol, const double *sigma9,
                                double *sigma_rot9);
int pfdd_host_resolss_raw(int ns, const double *xn, const double *xb, const double *sigma_rot9, double dslip,
                          double *tau);
int pfdd_host_resolss_non_schmid_raw(int ns, const double *xb, const double *sigma_rot9, double dslip, double angle,
                                     double w1, double w2, double w3, double *tau);
double pfdd_host_usfe_angle_tau_raw(const pfdd_host_material *m, const double *par5, const double *sigma_rot9,
                                    double angle);
int pfdd_host_project_core_energy(int nx, int ny, int nz, int ns, const double *xb, const double *xi, double *delta,
                                  double *ddelta);

#ifdef __cplusplus
}
#endif

#endif /* PFDD_HOST_H */
