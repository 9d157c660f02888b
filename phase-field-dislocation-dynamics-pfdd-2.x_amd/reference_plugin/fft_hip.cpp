/* fft_hip.cpp -- `fft_style hip`: the FFT engine of PFDD-2.x on one MI355X.
 *
 * Every pure virtual of the reference's FFT (src/fft.h:166-229) is
 * implemented.  The host keeps the arrays the rest of the reference reads
 * (xi, xo, xn, xb, tau, sigma_rot, data_eps, data_sigma, ave_*, theta, delta,
 * ddelta); the order parameters live on the GPU between output steps:
 *
 *   setup                 pfdd_gl_create (B(k) on the device) + pfdd_gl_set_xi
 *   rotate_stress, resolSS_*
 *                         host formulas (pfdd_host_*_raw), then pfdd_gl_set_tau
 *   init_loop .. prep_backward
 *                         nothing: core energy, the three forward FFTs, the
 *                         Green's-function contraction and the inverse FFTs
 *                         are one launch sequence inside update_order_parameter
 *   update_order_parameter
 *                         pfdd_gl_step(h, 1, 0, &r): xinorm, xiave, E_core;
 *                         xi[0] (and theta) are copied back when SolveGL is
 *                         about to call Output::compute (solve_GL.cpp:104-112)
 *   strain, stress        pfdd_gl_diag_fields -> data_eps, data_sigma, ave_eps,
 *                         ave_sigma (strain from the pre-update field, quirk
 *                         Q2 of the reference, as FFTW_Slab does)
 *   GLCond                pfdd_gl_cond_step once per iteration (in the first
 *                         backwardcond); the second backwardcond of
 *                         SolveGLCond::iterate reports conductivity_repeat
 *   initial_sxtal_*       pfdd_host_init_sxtal (cubic grids, scale 1 1 1)
 *
 * Errors of the C ABI abort through error->all, as the reference does.
 */
#include "fft_hip.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "app.h"
#include "error.h"
#include "material.h"
#include "memory.h"
#include "output.h"
#include "solve.h"
#include "pfdd_host.h"

#include <type_traits>

using namespace PFDD_NS;

FFT_HIP::FFT_HIP(PFDD_C *pfdd_p, int narg, char **arg)
    : FFT(pfdd_p, narg, arg), h(nullptr), hc(nullptr), cond_calls(0), iters_in_state(0), host_current(true),
      dslip_(0), bmag_(0), mu_s_(0)
{
  user_procgrid[0] = nprocs;
  user_procgrid[1] = user_procgrid[2] = 1;
  if (nprocs != 1)
    error->all(FLERR, "fft_style hip holds the whole grid on one GPU: run with one MPI rank");
  std::memset(&cr, 0, sizeof(cr));
  for (int q = 0; q < 9; q++) sr9[q] = 0.0;
}

FFT_HIP::~FFT_HIP()
{
  if (hc) pfdd_gl_cond_destroy(hc);
  if (h) pfdd_gl_destroy(h);
}

void FFT_HIP::check(int rc)
{
  if (rc) error->all(FLERR, pfdd_gl_last_error());
}

/* ---------------------------------------------------------------- set-up */

void FFT_HIP::create_plan() {}

void FFT_HIP::init()
{
  stress_inc = app->stoptime;
  dimension = app->dimension;
  sigma = app->sigma;
  deltasig = app->deltasig;
  create_plan();
  // one rank owns every x row
  local_x = nx;
  local_x_start = 0;
  local_ny_after_trans = ny;
  local_y_start_after_trans = 0;
  total_local_size = nx * ny * nz;
  local_y = ny;
  local_z = nz;
  allocate();
  // extent of the box along the normalised reciprocal axes (the frame the
  // initial configurations are laid out in): range over the eight corners,
  // starting from the origin
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int c = 0; c < 8; c++) {
    const double p[3] = {c & 1 ? boxxhi : boxxlo, c & 2 ? boxyhi : boxylo, c & 4 ? boxzhi : boxzlo};
    for (int a = 0; a < 3; a++) {
      const double t = (p[0] * four[a][0] + p[1] * four[a][1] + p[2] * four[a][2]) / normfour[a];
      if (t < lo[a]) lo[a] = t;
      if (t > hi[a]) hi[a] = t;
    }
  }
  boxxlof = lo[0]; boxxhif = hi[0];
  boxylof = lo[1]; boxyhif = hi[1];
  boxzlof = lo[2]; boxzhif = hi[2];
  for (int a = 0; a < 3; a++) cntf[a] = (hi[a] + lo[a]) / 2.0;
  xprdf = std::fabs(hi[0] - lo[0]);
  yprdf = std::fabs(hi[1] - lo[1]);
  zprdf = std::fabs(hi[2] - lo[2]);
}

void FFT_HIP::allocate()
{
  slip_systems = app->slip_systems;
  num_planes = app->num_planes;
  dimension = app->dimension;
  const int N = nx * ny * nz, S = slip_systems, ND = dimension;
  const int NP = num_planes > 0 ? num_planes : 1;
  const int NT = S > NP ? S : NP; // theta holds one angle per system and cell
  memory->create(xi, norder, 2 * S * N, "xi");
  memory->create(condxi, condnorder, 2 * N, "condxi");
  memory->create(xi_sum, norder, 2 * N, "xi_sum");
  for (int o = 0; o < norder; o++) {
    for (int q = 0; q < 2 * S * N; q++) xi[o][q] = 0.0;
    for (int q = 0; q < 2 * N; q++) xi_sum[o][q] = 0.0;
  }
  for (int o = 0; o < condnorder; o++)
    for (int q = 0; q < 2 * N; q++) condxi[o][q] = 0.0;
  memory->create(xo, S * N, "xo");
  for (int q = 0; q < S * N; q++) xo[q] = 0.0;
  memory->create(xn, S, ND, "xn");
  memory->create(xb, S, ND, "xb");
  memory->create(tau, S, "tau");
  for (int a = 0; a < S; a++) tau[a] = 0.0;
  memory->create(data_sigma, 2 * ND * ND * N, "data_sigma");
  memory->create(data_eps, 2 * ND * ND * N, "data_eps");
  memory->create(data_epsd, 2 * ND * ND * N, "data_epsd");
  for (int q = 0; q < 2 * ND * ND * N; q++) data_sigma[q] = data_eps[q] = data_epsd[q] = 0.0;
  memory->create(sigma_rot, ND, ND, "sigma_rot");
  memory->create(local_sigma, ND, ND, "local_sigma");
  memory->create(ave_sigma, ND, ND, "ave_sigma");
  memory->create(eps, S, ND, ND, "eps");
  memory->create(avepsd, ND, ND, "avepsd");
  memory->create(ave_epsd, ND, ND, "ave_epsd");
  memory->create(ave_eps, ND, ND, "ave_eps");
  memory->create(avepst, ND, ND, "avepst");
  memory->create(avepsts, ND, ND, "avepsts");
  memory->create(aveps, ND, ND, "aveps");
  double **m3[] = {sigma_rot, local_sigma, ave_sigma, avepsd, ave_epsd, ave_eps, avepst, avepsts, aveps};
  for (double **m : m3)
    for (int i = 0; i < ND; i++)
      for (int j = 0; j < ND; j++) m[i][j] = 0.0;
  memory->create(theta, NT * N, "theta");
  memory->create(delta, NP * N, "delta");
  memory->create(ddelta, NP * N, "ddelta");
  for (int q = 0; q < NT * N; q++) theta[q] = 0.0;
  for (int q = 0; q < NP * N; q++) delta[q] = ddelta[q] = 0.0;
  // FFTW work arrays, the host Green's functions (B, F, D) and the per-cell
  // core-energy scratch of FFTW_Slab have no host counterpart here
  BB = FF = DD = nullptr;
  fx = fy = fz = f = r = nullptr;
  gradx = grady = gradz = nullptr;
  fcore = f_core = df1core = df2core = df3core = dE_core = nullptr;
  C = nullptr;
}

void FFT_HIP::setup()
{
  rotate_stress();
  frec();
  stiffness();
  greens_function();
  Bmatrix();
  Fmatrix();
  if (!h) create_handle();
  else if (!host_current) sync_host();
  upload_state();
}

void FFT_HIP::frec() {}            // frequencies are formed on the device (x pass)
void FFT_HIP::stiffness() {}       // isotropic C from C44, C12 on the device
void FFT_HIP::greens_function() {} // Green's function evaluated per k-point on the device
void FFT_HIP::Fmatrix() {}         // strain operator applied on the device (pfdd_gl_diag_fields)

// The material's Voigt moduli, as Bmatrix stores them back into the
// material (fft_fftw_slab.cpp:487-496); B(k) itself is built on the device.
void FFT_HIP::Bmatrix()
{
  const double d = (2.0 * material->C44 + material->C12 - material->C11) / 5.0;
  const double mu = material->C44 - d, ll = material->C12 - d;
  const double young = mu * (3 * ll + 2 * mu) / (ll + mu);
  material->mu = mu;
  material->ll = ll;
  material->young = young;
  material->xnu = young / 2.0 / mu - 1.0;
  mu_s_ = mu;
  dslip_ = material->dslip;
  bmag_ = material->b;
}

void FFT_HIP::create_handle()
{
  const int S = slip_systems, N = nx * ny * nz;
  const double mdb = mu_s_ * dslip_ * bmag_;
  pfdd_gl_desc d;
  std::memset(&d, 0, sizeof(d));
  d.nx = nx;
  d.ny = ny;
  d.nz = nz;
  d.ns = S;
  d.precision = PFDD_FP64;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) d.four[3 * i + j] = four[i][j];
  d.C44 = material->C44;
  d.C12 = material->C12;
  d.C11 = material->C11;
  d.S12 = material->S12;
  d.S44 = material->S44;
  d.dslip = dslip_;
  d.xn = &xn[0][0];
  d.xb = &xb[0][0];
  d.cddt = app->CD * app->timestep;
  // core_energy_* normalisations (fft_fftw_slab.cpp:1090-1094, 1165-1178,
  // 1281-1287, 1396-1406, 1468-1476, 1231-1232)
  std::vector<double> coef, amp;
  switch (app->core_flag) {
  case 1:
    d.core_model = PFDD_CORE_PERFECT;
    coef = {material->An / mdb};
    break;
  case 2:
    d.core_model = PFDD_CORE_EXTENDED;
    coef = {material->c0 / mdb, material->c1 / mdb, material->c2 / mdb, material->c3 / mdb, material->c4 / mdb,
            material->a1 / mdb, material->a3 / mdb};
    break;
  case 3: {
    d.core_model = PFDD_CORE_SINE;
    const double isf = material->isf / mdb;
    coef = {isf, (material->usf - isf / 2.0) / mdb};
    break;
  }
  case 4:
    d.core_model = PFDD_CORE_PYRII;
    coef = {material->aa0 / mdb, material->aa1 / mdb, material->aa2 / mdb, material->aa3 / mdb,
            material->aa4 / mdb, material->bb1 / mdb, material->bb2 / mdb, material->bb3 / mdb,
            material->bb4 / mdb};
    break;
  case 5:
    d.core_model = PFDD_CORE_BCC;
    coef = {material->usf / mdb};
    break;
  case 6: // An'(angle, tau) refreshed in core_energy_USFE_angle_tau
    d.core_model = PFDD_CORE_SERIES;
    d.ncoef = 1;
    coef.assign(S, 0.0);
    break;
  case 7:
    if (!material->A_mpea) error->all(FLERR, "core flag 7 needs the material's MPEA file");
    d.core_model = PFDD_CORE_MPEA;
    amp.resize((size_t)S * N);
    for (size_t q = 0; q < amp.size(); q++) amp[q] = material->A_mpea[q] / mdb;
    d.amp = amp.data();
    break;
  default:
    d.core_model = PFDD_CORE_NONE;
  }
  d.coef = coef.empty() ? nullptr : coef.data();
  d.xo = xo;
  d.device = 0;
  d.flags = PFDD_GL_FLAG_AUTOTUNE;
  check(pfdd_gl_create(&d, &h));
}

void FFT_HIP::upload_state()
{
  check(pfdd_gl_set_xi(h, xi[0]));
  check(pfdd_gl_set_tau(h, tau));
  host_current = true;
}

// Field (and character angle) back to the host arrays Output and the dumps read.
void FFT_HIP::sync_host()
{
  if (host_current || !h) return;
  check(pfdd_gl_get_xi(h, xi[0]));
  check(pfdd_gl_theta(h, theta));
  host_current = true;
}

// SolveGL::iterate calls Output::compute after this iteration
// (solve_GL.cpp:104-112; xinorm < tol, the last iteration of a stress state,
// or a scheduled output time; ximin stays 0).
bool FFT_HIP::output_due() const
{
  return xinorm < solve->tol || iters_in_state >= solve->max_iter || app->time >= app->nextoutput;
}

/* ---------------------------------------------------------- stress state */

void FFT_HIP::rotate_stress()
{
  double s9[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) s9[3 * i + j] = sigma[i][j];
  if (pfdd_host_rotate_stress_raw(&prim[0][0], &invprim[0][0], norvol, s9, sr9))
    error->all(FLERR, "rotate_stress: invalid geometry");
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) sigma_rot[i][j] = sr9[3 * i + j];
  iters_in_state = 0;
}

void FFT_HIP::resolSS_Schmid()
{
  if (pfdd_host_resolss_raw(slip_systems, &xn[0][0], &xb[0][0], sr9, material->dslip, tau))
    error->all(FLERR, "resolSS_Schmid: invalid slip systems");
  if (h) check(pfdd_gl_set_tau(h, tau));
}

void FFT_HIP::resolSS_non_Schmid()
{
  if (pfdd_host_resolss_non_schmid_raw(slip_systems, &xb[0][0], sr9, material->dslip, app->angle_to_110, app->w1,
                                       app->w2, app->w3, tau))
    error->all(FLERR, "resolSS_non_Schmid: invalid slip systems");
  if (h) check(pfdd_gl_set_tau(h, tau));
}

/* --------------------------------------------------------- GL iteration */

void FFT_HIP::init_loop()
{
  // the projected order parameter is cleared every iteration (init_loop of
  // the reference); project_core_energy refills it after the output step
  if (app->non_Schmid) {
    const int n = (num_planes > 0 ? num_planes : 1) * nx * ny * nz;
    for (int q = 0; q < n; q++) delta[q] = ddelta[q] = 0.0;
  }
}

void FFT_HIP::gradient() {}
void FFT_HIP::core_energy_bcc_perfect() {}
void FFT_HIP::core_energy_perfect() {}
void FFT_HIP::core_energy_extended() {}
void FFT_HIP::core_energy_sine() {}
void FFT_HIP::core_energy_pyrII() {}
void FFT_HIP::core_energy_mpea() {}

// USFE(angle, tau) of the current stress state as the one coefficient of a
// sin^2 core (core_energy_USFE_angle_tau, fft_fftw_slab.cpp:1500-1545)
void FFT_HIP::core_energy_USFE_angle_tau()
{
  pfdd_host_material m;
  std::memset(&m, 0, sizeof(m));
  m.mu_s = mu_s_;
  m.dslip = dslip_;
  m.b = bmag_;
  const double par[5] = {material->a_slope, material->a_b, material->b_b, material->c_slope, material->c_b};
  const double An = pfdd_host_usfe_angle_tau_raw(&m, par, sr9, app->angle_to_110);
  std::vector<double> c(slip_systems, An);
  check(pfdd_gl_set_core_coef(h, 1, c.data()));
}

void FFT_HIP::prep_forward() {}
void FFT_HIP::forward_mode1() {}
void FFT_HIP::forward_mode2() {}
void FFT_HIP::internal_energy() {}
void FFT_HIP::prep_backward() {}
void FFT_HIP::backward_mode1() {}
void FFT_HIP::backward_mode2() {}

void FFT_HIP::update_order_parameter()
{
  pfdd_gl_reduce red;
  check(pfdd_gl_step(h, 1, 0.0, &red));
  xinorm = red.xinorm;
  xiave = red.xiave;
  E_core = red.e_core_ref;
  ximin = 0.0;
  host_current = false;
  iters_in_state++;
  if (output_due()) sync_host();
}

void FFT_HIP::prepare_next_itr() {}

void FFT_HIP::project_core_energy()
{
  sync_host();
  if (pfdd_host_project_core_energy(nx, ny, nz, slip_systems, &xb[0][0], xi[0], delta, ddelta))
    error->all(FLERR, "project_core_energy: invalid grid");
}

void FFT_HIP::temp_stor_data() {}
void FFT_HIP::reset_data() {}

/* ---------------------------------------------------------- output step */

// strain(), total_average_strain() and stress() of the reference in one
// device call; total_average_strain and stress reuse its results
void FFT_HIP::strain()
{
  double ae[9], as[9];
  check(pfdd_gl_diag_fields(h, sr9, ae, as, data_eps, data_sigma));
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      ave_eps[i][j] = ae[3 * i + j];
      ave_sigma[i][j] = as[3 * i + j];
    }
}

void FFT_HIP::stressfree_strain() {}
void FFT_HIP::total_average_strain() {}
void FFT_HIP::stress() {}

/* --------------------------------------------------------- conductivity */

void FFT_HIP::forwardcond()
{
  if (!hc) {
    // per-row layer conductivities of backwardcond's two-material stack
    Material *A = materialA ? materialA : material, *B = materialB ? materialB : A;
    auto params = [](const Material *m, double *p) {
      const double v[11] = {m->condsigma, m->mfpel, m->alpha_resistivity, m->alpha_mfpel, m->Rgrain,
                            m->beta_disloc, m->rho_disloc, m->dlayer, (double)m->Nlayers, (double)m->Ngrains,
                            m->pinterface};
      for (int q = 0; q < 11; q++) p[q] = v[q];
    };
    double pa[11], pb[11];
    params(A, pa);
    params(B, pb);
    std::vector<double> sg(ny), ds(ny);
    if (pfdd_host_cond_layers(pa, pb, ny, app->temperature, sg.data(), ds.data()))
      error->all(FLERR, "conductivity: invalid layer parameters");
    pfdd_gl_cond_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) d.four[3 * i + j] = four[i][j];
    d.sigma = sg.data();
    d.dsigma = ds.data();
    d.El = material->El[1];
    d.epsilonzero = material->epsilonzero;
    d.rate = 1e-22 * app->timestep;
    d.device = 0;
    check(pfdd_gl_cond_create(&d, &hc));
  }
  cond_calls = 0;
}

void FFT_HIP::divcurrent() {}

// SolveGLCond::iterate calls backwardcond twice per iteration
// (solve_GL+cond.cpp:107,137): the first advances the conductivity field, the
// second reports backwardcond applied to its own real-space output
void FFT_HIP::backwardcond()
{
  if (cond_calls == 0) {
    check(pfdd_gl_cond_step(hc, 1, &cr));
    conductivity = cr.conductivity;
  } else {
    conductivity = cr.conductivity_repeat;
  }
  conductivity_global = conductivity;
  cond_calls++;
}

void FFT_HIP::update_conduct_order_param()
{
  condxinorm = cr.condxinorm;
  condxiave = cr.condxiave;
}

void FFT_HIP::prepare_next_conditr() {}

/* ------------------------------------------------ initial configurations */

void FFT_HIP::init_style(const char *style)
{
  if (pfdd_host_init_sxtal(style, nx, ny, nz, slip_systems, xi[0], xo)) {
    char msg[160];
    std::snprintf(msg, sizeof(msg), "initial configuration %s: not available for this grid in fft_style hip", style);
    error->all(FLERR, msg);
  }
  host_current = true;
  if (h) upload_state();
}

void FFT_HIP::initial_sxtal_mode1() { init_style("mode1"); }
void FFT_HIP::initial_sxtal_mode2() { init_style("mode2"); }
void FFT_HIP::initial_sxtal_NoLoop() { init_style("NoLoop"); }
void FFT_HIP::initial_sxtal_2LOrtho() { init_style("2LOrtho"); }
void FFT_HIP::initial_sxtal_1LOrthoFCC() { init_style("1LOrthoFCC"); }
void FFT_HIP::initial_sxtal_1LNonOrthoFCC() { init_style("1LNonOrthoFCC"); }
void FFT_HIP::initial_sxtal_1LInclinedFCC() { init_style("1LInclinedFCC"); }
void FFT_HIP::initial_sxtal_1LOrthoBCC() { init_style("1LOrthoBCC"); }
void FFT_HIP::initial_sxtal_StraightOrtho() { init_style("StraightOrtho"); }
void FFT_HIP::initial_sxtal_StraightNonOrtho() { init_style("StraightNonOrtho"); }
void FFT_HIP::initial_sxtal_3D3ShcpBasal() { init_style("3D3ShcpBasal"); }
void FFT_HIP::initial_sxtal_3D3ShcpNotch() { init_style("3D3ShcpNotch"); }
void FFT_HIP::initial_sxtal_frank_read() { init_style("frank_read"); }
void FFT_HIP::initial_sxtal_3SBCC() { init_style("3SBCC"); }
void FFT_HIP::initial_sxtal_edge() { init_style("edge"); }
void FFT_HIP::initial_sxtal_loop() { init_style("loop"); }

/* ----------------------------------------------------------- exchange */

void FFT_HIP::fft_send(void *buf, int count, MPI_Datatype type, int dest, int tag, MPI_Comm comm)
{
  MPI_Send(buf, count, type, dest, tag, comm);
}

void FFT_HIP::fft_recv(void *buf, int count, MPI_Datatype type, int src, int tag, MPI_Comm comm, MPI_Status *st)
{
  MPI_Recv(buf, count, type, src, tag, comm, st);
}

// every pure virtual of FFT (src/fft.h:166-229) has its implementation above
static_assert(!std::is_abstract<FFT_HIP>::value, "FFT_HIP must implement every pure virtual of FFT");
