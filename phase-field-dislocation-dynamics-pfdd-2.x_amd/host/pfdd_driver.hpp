// pfdd_driver.hpp -- host-side mirror of the PFDD-2.x classes that drive the GL
// path (FFT, SolveGL, App, Output, Input), running the per-step work on the
// GPU through the C ABI of include/pfdd_gl.h.
//
//   FFT_HIP   the FFT interface of fft.h:24-231 restricted to what SolveGL,
//             App and Output call on the GL path; every per-step virtual maps
//             onto one fused GPU iteration (pfdd_gl_step)
//   SolveGL   SolveGL::iterate (solve_GL.cpp:70-158): stress states x GL
//             iterations, output trigger, xinorm < tol break, sigma ramp
//   App       app_style slip tables, CD/timestep, initial field, run/stoptime
//   Output    stats line (output.cpp:325-355), strain/stress diagnostics
//             (diag_strain.cpp:44-68, diag_stress.cpp:43-80), text dumps
//   Input     the in.* deck grammar (input.cpp:117-190, 438-495)
#pragma once

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pfdd_gl.h"
#include "pfdd_host.hpp"

namespace pfdd_host {

struct DeckError {
    std::string msg;
};

class FFT_HIP {
  public:
    // state mirrored from class FFT (fft.h:24-150)
    int nx = 0, ny = 0, nz = 0, slip_systems = 0, num_planes = 0, dimension = 3, mode = 1;
    int precision = PFDD_FP64, flags = 0;
    double xinorm = 0, ximin = 0, xiave = 0, E_core = 0;
    std::vector<double> xi, xo; // host mirrors of xi[0] and xo
    double sigma[3][3] = {}, deltasig[3][3] = {}, sigma_rot[3][3] = {};
    std::vector<double> tau;
    double ave_eps[3][3] = {}, ave_sigma[3][3] = {};
    std::vector<double> data_eps, data_sigma; // fft->data_eps / data_sigma (2*9*N), when dumped
    bool want_fields = false;
    Geometry geo;
    Material mat;
    SlipTable slip;
    int core_flag = 1;
    std::vector<double> coef;
    int ncoef = 0;
    double CD = 0.5, timestep = 1.0;
    pfdd_gl_reduce last{};
    bool host_current = true;

    ~FFT_HIP();
    void init();  // FFTW_Slab::init + allocate + setup: device handle, B(k)
    void rotate_stress();
    void resolSS_Schmid();
    void resolSS_non_Schmid();
    // App::resolSS (app.cpp:654-658) + the per-stress-state An' of core flag 6
    // (core_energy_USFE_angle_tau, fft_fftw_slab.cpp:1500-1545)
    void resolSS();
    bool non_schmid = false;                      // non_Schmid command (app.cpp:446-460)
    double angle_to_110 = 0, w1 = 0, w2 = 0, w3 = 0;
    // The per-iteration virtuals of SolveGL::iterate.  init_loop, gradient,
    // core_energy, prep_forward, internal_energy, prep_backward and
    // prepare_next_itr are fused into the GPU iteration that
    // update_order_parameter launches.
    void init_loop() {}
    void gradient() {}
    void core_energy() {}
    void prep_forward() {}
    void internal_energy() {}
    void prep_backward() {}
    void update_order_parameter() { iterate_block(1, 0.0); }
    void prepare_next_itr() {}
    // n iterations in one call; stops on the device after the first one with
    // xinorm < tol.  Returns the number executed.
    int iterate_block(int n, double tol);
    void diagnostics();   // strain() + total_average_strain() + stress(): averages (+ fields)
    void sync_host();     // xi device -> host mirror
    void upload_xi();     // host mirror -> device
    pfdd_gl *handle() const { return h_; }

    // conductivity sub-problem of solve_style GLCond: the second material
    // command (fft->materialB), the temperature command, the per-row layer
    // conductivities, and the GPU field c (fft->condxi)
    // fft->delta / ddelta of plane 0 (project_core_energy, fft_fftw_slab.cpp:
    // 1550-1624), the arrays the dump fields deltao / ddelta read
    std::vector<double> delta, ddelta;
    void project_core_energy();
    // fft->theta (gradient(), computed at the start of each iteration); zero
    // before the first iteration, as the freshly allocated array reads
    bool stepped = false;
    void theta(std::vector<double> &th);
    Material matB;
    bool have_matB = false, cond = false;
    double temperature = 300.0;
    std::vector<double> cond_sigma, cond_dsigma;
    pfdd_gl_cond_reduce cond_last{};
    void cond_init();
    // forwardcond + divcurrent + backwardcond + update_conduct_order_param
    // (+ prepare_next_conditr) of one iteration
    void cond_iterate();
    void cond_get(std::vector<double> &c);

  private:
    pfdd_gl *h_ = nullptr;
    pfdd_gl_cond *c_ = nullptr;
};

struct DumpSpec {
    std::string id, file;
    double every = 1;
    std::vector<std::string> fields;
    double next_time = 0;
    int idump = 0;       // snapshots written (Dump::idump, dump.cpp:114,172)
    bool binary = false; // file name ends in .bin (dump.cpp:81-82)
};

class Output {
  public:
    double stats_delta = 0.0, stats_time = 0.0;
    bool diag_strain = false, diag_stress = false;
    std::vector<DumpSpec> dumps;
    FILE *screen = stdout, *logfile = nullptr;
    double loop_time = 0.0;
    double setup(class App &app, FFT_HIP &fft, double time);
    double compute(class App &app, FFT_HIP &fft, double time, int done);
    void stats_header();
    void stats(App &app, FFT_HIP &fft, int timeflag);
    void write_dump(DumpSpec &d, App &app, FFT_HIP &fft, double time);
    static double next_time(double tcurrent, double delta);
};

class App {
  public:
    std::string style;
    int core_flag = -1;
    double time = 0.0, stoptime = 0.0, nextoutput = 0.0;
    int stopsteps = 0, steps = 0;
    double CD = -1, timestep = -1; // < 0: app defaults
    double scale[3] = {1, 1, 1};
    bool scale_set = false;
    double sigma0[3][3] = {}, dsigma[3][3] = {};
    // initial field
    std::string init_style = "loops";
    bool init_set = false; // an initial command was given (else the app's default setup)
    int init_nloops = 0, seed = 576;
    double init_rmin = 8, init_rmax = 20, init_noise = 0.0;
    struct Disc {
        int a, axis, plane;
        double c0, c1, r;
    };
    std::vector<Disc> discs;
    void init_app(FFT_HIP &fft);
};

class SolveGL {
  public:
    int max_iter = 10;
    double tol = 0.0;
    long long gpu_calls = 0;
    bool cond = false; // solve_style GLCond
    void iterate(App &app, FFT_HIP &fft, Output &out);
    // SolveGLCond::iterate (solve_GL+cond.cpp:70-150) as executed
    void iterate_cond(App &app, FFT_HIP &fft, Output &out);
};

// Parses and runs an in.* deck.  Throws DeckError on a malformed deck (the
// reference's Error::all).  check_only: stop before touching the GPU and
// print the derived setup.
class Input {
  public:
    FFT_HIP fft;
    App app;
    SolveGL solve;
    Output output;
    bool box = false, sites = false, have_material = false, check_only = false;
    std::string write_xi_file;
    int run_count = 0;
    void file(FILE *fp);
    void line(const std::string &text);
    void command(const std::vector<std::string> &w);
    void run(const std::vector<std::string> &w);
    void print_setup();
};

} // namespace pfdd_host
