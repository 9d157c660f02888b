/* fft_hip.h -- `fft_style hip` for PFDD-2.x: the FFT engine of the reference
 * (the abstract class FFT of src/fft.h:24-231) implemented on one MI355X
 * through this repository's C ABI (include/pfdd_gl.h, include/pfdd_host.h).
 *
 * Copy this header and fft_hip.cpp into the reference's src/ directory, add
 * -I<repo>/include to CCFLAGS and
 *   -L<repo>/phase-field-dislocation-dynamics-pfdd-2.x_amd/lib -lpfdd_gl -lpfdd_host
 *   -L/opt/rocm/lib -lamdhip64
 * to LINKFLAGS; Make.csh collects the FftStyle line below into style_fft.h
 * like every other style (fft_fftw_slab.h:16-17), and a deck selects it with
 * `fft_style hip ...` (input.cpp:693-708).  SolveGL, SolveGLCond, the apps,
 * Output, the diagnostics and the dumps are unchanged.
 *
 * One MPI rank drives one GPU: the whole grid lives on the device, so the
 * slab decomposition of FFTW_Slab collapses to local_x = nx.
 */
#ifdef FFT_CLASS
FftStyle(hip,FFT_HIP)

#else

#ifndef PFDD_FFT_HIP_H
#define PFDD_FFT_HIP_H

#include "mpi.h"
#include "fft.h"
#include "pfdd_gl.h"

namespace PFDD_NS {

class FFT_HIP : public FFT {
 public:
  FFT_HIP(class PFDD_C *, int, char **);
  ~FFT_HIP();

  // set-up: host mirrors of the FFT arrays, GPU handle, B(k) on the device
  void create_plan();
  void init();
  void setup();
  void allocate();
  void frec();
  void stiffness();
  void greens_function();
  void Bmatrix();
  void Fmatrix();

  // stress state: host formulas, tau uploaded to the device
  void rotate_stress();
  void resolSS_Schmid();
  void resolSS_non_Schmid();

  // one GL iteration: everything up to the update is fused into one launch
  // sequence (pfdd_gl_step), so these record nothing
  void init_loop();
  void gradient();
  void core_energy_bcc_perfect();
  void core_energy_perfect();
  void core_energy_extended();
  void core_energy_sine();
  void core_energy_pyrII();
  void core_energy_USFE_angle_tau();
  void core_energy_mpea();
  void prep_forward();
  void forward_mode1();
  void forward_mode2();
  void internal_energy();
  void prep_backward();
  void backward_mode1();
  void backward_mode2();
  void update_order_parameter();
  void prepare_next_itr();
  void project_core_energy();
  void temp_stor_data();
  void reset_data();

  // output steps (Output::compute, DiagStrain, DiagStress, DumpText)
  void stressfree_strain();
  void total_average_strain();
  void strain();
  void stress();

  // conductivity (solve_style GLCond)
  void forwardcond();
  void divcurrent();
  void backwardcond();
  void update_conduct_order_param();
  void prepare_next_conditr();

  // initial configurations (App::set_initial_sxtal, app.cpp:810-885)
  void initial_sxtal_mode1();
  void initial_sxtal_mode2();
  void initial_sxtal_NoLoop();
  void initial_sxtal_2LOrtho();
  void initial_sxtal_1LOrthoFCC();
  void initial_sxtal_1LNonOrthoFCC();
  void initial_sxtal_1LInclinedFCC();
  void initial_sxtal_1LOrthoBCC();
  void initial_sxtal_StraightOrtho();
  void initial_sxtal_StraightNonOrtho();
  void initial_sxtal_3D3ShcpBasal();
  void initial_sxtal_3D3ShcpNotch();
  void initial_sxtal_frank_read();
  void initial_sxtal_3SBCC();
  void initial_sxtal_edge();
  void initial_sxtal_loop();

  // point-to-point exchange of FFTW_Slab's halo code; one rank here
  void fft_send(void *, int, MPI_Datatype, int, int, MPI_Comm);
  void fft_recv(void *, int, MPI_Datatype, int, int, MPI_Comm, MPI_Status *);

 private:
  pfdd_gl *h;               // GL solver on the device
  pfdd_gl_cond *hc;         // conductivity field on the device (GLCond runs)
  pfdd_gl_cond_reduce cr;   // scalars of the last conductivity iteration
  int cond_calls;           // backwardcond calls since forwardcond (GLCond calls it twice)
  int iters_in_state;       // GL iterations since rotate_stress (a new stress state)
  bool host_current;        // xi[0] holds the device field
  double sr9[9];            // sigma_rot, row-major (units of mu)
  double dslip_, bmag_, mu_s_;

  void check(int rc);
  void create_handle();
  void upload_state();
  void sync_host();
  bool output_due() const;
  void init_style(const char *style);
};

}

#endif
#endif
