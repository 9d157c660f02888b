// pfdd_driver.cpp -- see pfdd_driver.hpp.
#include "pfdd_driver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace pfdd_host {

namespace {

[[noreturn]] void deck_error(const std::string &m) { throw DeckError{m}; }

double wall()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

double num(const std::vector<std::string> &w, size_t i, const char *what)
{
    if (i >= w.size()) deck_error(std::string("Illegal ") + what + " command");
    char *end = nullptr;
    const double v = std::strtod(w[i].c_str(), &end);
    if (end == w[i].c_str() || *end) deck_error(std::string("Illegal ") + what + " command: '" + w[i] + "'");
    return v;
}

void check(int rc)
{
    if (rc != PFDD_OK) deck_error(std::string("GPU solver: ") + pfdd_gl_last_error());
}

} // namespace

// ===================================================================== FFT_HIP
FFT_HIP::~FFT_HIP()
{
    if (c_) pfdd_gl_cond_destroy(c_);
    if (h_) pfdd_gl_destroy(h_);
}

void FFT_HIP::project_core_energy()
{
    // project_core_energy (fft_fftw_slab.cpp:1550-1624) for plane 0, the one
    // the apps' compute_delta / compute_ddelta read: delta = Re xi_0 for one or
    // two systems, else the projection of the plane's three systems onto b_1;
    // ddelta = forward difference along x on the line j = ny/2, k = nz/2
    // (dx = 1), zero at the last cell and off the line
    sync_host();
    const long long N = (long long)nx * ny * nz;
    delta.assign(N, 0.0);
    ddelta.assign(N, 0.0);
    pfdd_host::project_core_energy(nx, ny, nz, slip_systems, slip.xb.data(), xi.data(), delta.data(), ddelta.data());
}

void FFT_HIP::theta(std::vector<double> &th)
{
    th.assign((size_t)slip_systems * nx * ny * nz, 0.0);
    if (stepped) check(pfdd_gl_theta(h_, th.data()));
}

void FFT_HIP::cond_init()
{
    // FFTW_Slab::allocate's condxi (zero) and backwardcond's layer algebra,
    // which needs two materials (fft_fftw_slab.cpp:2083)
    if (!have_matB) deck_error("Need two materials for this app!");
    std::string msg;
    if (!cond_layers(mat, matB, ny, temperature, cond_sigma, cond_dsigma, &msg)) deck_error(msg);
    pfdd_gl_cond_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d.four[3 * i + j] = geo.four[i][j];
    d.sigma = cond_sigma.data();
    d.dsigma = cond_dsigma.data();
    d.El = mat.El[1];
    d.epsilonzero = mat.epsilonzero;
    d.rate = 1e-22 * timestep;
    d.device = 0;
    check(pfdd_gl_cond_create(&d, &c_));
    const double Vf = matB.dlayer / (mat.dlayer + matB.dlayer);
    printf("temperature: %f \n", temperature);
    printf("size in y: %i Nmaterials: %i \n", ny, 2);
    printf("materialA: conductivity = %f MS/m, mean-free-path = %f nm, dlayer = %f nm\n", mat.condsigma / 1e6,
           1.e9 * mat.mfpel, 1.e9 * mat.dlayer);
    printf("materialB: conductivity = %f MS/m, mean-free-path = %f nm, dlayer = %f nm\n", matB.condsigma / 1e6,
           1.e9 * matB.mfpel, 1.e9 * matB.dlayer);
    printf("Nlayers: %i, Vf = %f \n", mat.Nlayers, Vf);
}

void FFT_HIP::cond_iterate()
{
    check(pfdd_gl_cond_step(c_, 1, &cond_last));
}

void FFT_HIP::cond_get(std::vector<double> &c)
{
    c.assign((size_t)2 * nx * ny * nz, 0.0);
    check(pfdd_gl_cond_get(c_, c.data()));
}

void FFT_HIP::init()
{
    pfdd_gl_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.ns = slip_systems;
    d.precision = precision;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d.four[3 * r + c] = geo.four[r][c];
    d.C44 = mat.C44;
    d.C12 = mat.C12;
    d.C11 = mat.C11;
    d.S12 = mat.S12;
    d.S44 = mat.S44;
    d.dslip = mat.dslip;
    d.xn = slip.xn.data();
    d.xb = slip.xb.data();
    d.cddt = CD * timestep;
    std::vector<double> amp;
    if (core_flag == 1) {
        d.core_model = PFDD_CORE_PERFECT;
    } else if (core_flag == 3) {
        d.core_model = PFDD_CORE_SINE;
    } else if (core_flag == 7) {
        d.core_model = PFDD_CORE_MPEA;
        const double mdb = mat.mu_s * mat.dslip * mat.b;
        amp.resize(mat.A_mpea.size());
        for (size_t q = 0; q < amp.size(); ++q) amp[q] = mat.A_mpea[q] / mdb;
        d.amp = amp.data();
    } else if (core_flag == 8) {
        d.core_model = PFDD_CORE_SERIES;
    } else if (core_flag == 6) {
        // sine core with An' of the stress state (set by resolSS each state)
        d.core_model = PFDD_CORE_PERFECT;
    } else if (core_flag == 5) {
        d.core_model = PFDD_CORE_BCC;
    } else if (core_flag == 2) {
        d.core_model = PFDD_CORE_EXTENDED;
    } else if (core_flag == 4) {
        d.core_model = PFDD_CORE_PYRII;
    } else if (core_flag == 9) {
        d.core_model = PFDD_CORE_FOURIER;
    } else if (core_flag == 0) {
        d.core_model = PFDD_CORE_NONE;
    } else {
        deck_error("core_energy " + std::to_string(core_flag) + " is not available on the GPU path");
    }
    d.ncoef = ncoef;
    d.coef = coef.empty() ? nullptr : coef.data();
    bool any = false;
    for (double v : xo) any |= v != 0.0;
    d.xo = any ? xo.data() : nullptr;
    d.device = 0;
    d.flags = flags;
    check(pfdd_gl_create(&d, &h_));
    upload_xi();
}

void FFT_HIP::rotate_stress()
{
    pfdd_host::rotate_stress(geo, sigma, sigma_rot);
}

void FFT_HIP::resolSS_Schmid()
{
    tau.assign(slip_systems, 0.0);
    resolss(slip, sigma_rot, mat.dslip, tau.data());
    check(pfdd_gl_set_tau(h_, tau.data()));
}

void FFT_HIP::resolSS_non_Schmid()
{
    tau.assign(slip_systems, 0.0);
    resolss_non_schmid(slip, sigma_rot, mat.dslip, angle_to_110, w1, w2, w3, tau.data());
    printf("angle %f, w1 %f,w2 %f,w3 %f \n", angle_to_110, w1, w2, w3);
    for (int is = 0; is < slip_systems; ++is) printf("nonSchmid: Resolved shear stresses tau[%d] = %f\n", is, tau[is]);
    check(pfdd_gl_set_tau(h_, tau.data()));
}

void FFT_HIP::resolSS()
{
    if (non_schmid) resolSS_non_Schmid();
    else resolSS_Schmid();
    if (core_flag == 6) {
        // An' depends on the stress state: refreshed once per state, it is
        // constant over the state's GL iterations
        const double An = usfe_angle_tau(mat, sigma_rot, angle_to_110);
        printf("core flag=6, usfe(angle,tau): %f\n", An * mat.mu_s * mat.dslip * mat.b * 1.0E3);
        std::vector<double> c(slip_systems, An);
        check(pfdd_gl_set_core_coef(h_, 1, c.data()));
    }
}

int FFT_HIP::iterate_block(int n, double tol)
{
    check(pfdd_gl_step(h_, n, tol, &last));
    stepped = true;
    xinorm = last.xinorm;
    xiave = last.xiave;
    E_core = last.e_core_ref;
    host_current = false;
    return last.steps_done;
}

void FFT_HIP::diagnostics()
{
    double e[9], s[9], sr[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) sr[3 * r + c] = sigma_rot[r][c];
    if (want_fields) {
        const size_t n = (size_t)18 * nx * ny * nz;
        if (data_eps.size() != n) data_eps.assign(n, 0.0);
        if (data_sigma.size() != n) data_sigma.assign(n, 0.0);
        check(pfdd_gl_diag_fields(h_, sr, e, s, data_eps.data(), data_sigma.data()));
    } else {
        check(pfdd_gl_diag(h_, sr, e, s));
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            ave_eps[r][c] = e[3 * r + c];
            ave_sigma[r][c] = s[3 * r + c];
        }
}

void FFT_HIP::sync_host()
{
    if (host_current) return;
    check(pfdd_gl_get_xi(h_, xi.data()));
    host_current = true;
}

void FFT_HIP::upload_xi()
{
    check(pfdd_gl_set_xi(h_, xi.data()));
    host_current = true;
}

// ====================================================================== Output
double Output::next_time(double tcurrent, double delta)
{
    // Output::next_time, logfreq 0 (output.cpp:393-400)
    double tnew = std::ceil(tcurrent / delta) * delta;
    if (tnew == tcurrent) tnew = tcurrent + delta;
    return tnew;
}

void Output::stats_header()
{
    std::string s;
    char buf[64];
    snprintf(buf, sizeof buf, " %10s %10s", "Time", "Step");
    s += buf;
    snprintf(buf, sizeof buf, " %10s", "CPU");
    s += buf;
    if (diag_strain) {
        const char *h[6] = {"e11", "e22", "e33", "e12", "e13", "e23"};
        for (auto x : h) {
            snprintf(buf, sizeof buf, " %8s", x);
            s += buf;
        }
    }
    if (diag_stress) {
        const char *h[6] = {"s11", "s22", "s33", "s12", "s13", "s23"};
        for (auto x : h) {
            snprintf(buf, sizeof buf, " %8s", x);
            s += buf;
        }
    }
    fprintf(screen, "%s\n", s.c_str());
    if (logfile) fprintf(logfile, "%s\n", s.c_str());
}

void Output::stats(App &app, FFT_HIP &fft, int timeflag)
{
    // Output::stats (output.cpp:325-355) with the app's time/steps columns and
    // the %8g diagnostics columns (diag_strain.cpp:64-68, diag_stress.cpp:76-80)
    std::string s;
    char buf[128];
    snprintf(buf, sizeof buf, " %10g %10d", app.time, app.steps);
    s += buf;
    snprintf(buf, sizeof buf, " %10.3g", timeflag ? loop_time : 0.0);
    s += buf;
    if (diag_strain) {
        const double *e = &fft.ave_eps[0][0];
        snprintf(buf, sizeof buf, " %8g %8g %8g %8g %8g %8g", e[0], e[4], e[8], e[1], e[2], e[5]);
        s += buf;
    }
    if (diag_stress) {
        const double *t = &fft.ave_sigma[0][0];
        snprintf(buf, sizeof buf, " %8g %8g %8g %8g %8g %8g", t[0], t[4], t[8], t[1], t[2], t[5]);
        s += buf;
    }
    fprintf(screen, "%s\n", s.c_str());
    if (logfile) {
        fprintf(logfile, "%s\n", s.c_str());
        fflush(logfile);
    }
}

void Output::write_dump(DumpSpec &d, App &app, FFT_HIP &fft, double time)
{
    // Dump::write (dump.cpp:145-233) for one process: the file is created at
    // the first snapshot and appended to afterwards.  Text: DumpText::header_text
    // (dump_text.cpp:361-372) + one line per site; binary (.bin):
    // DumpText::header_binary (:341-357) -- idump, time, ndump, box bounds in
    // lattice units, size_one, nprocs = 1 -- then write_binary (:376-381):
    // the count of doubles and the site-major array.
    fft.sync_host();
    FILE *fp = fopen(d.file.c_str(), d.idump == 0 ? (d.binary ? "wb" : "w") : (d.binary ? "ab" : "a"));
    if (!fp) deck_error("Cannot open dump file " + d.file);
    const long long N = (long long)fft.nx * fft.ny * fft.nz;
    const int size_one = (int)d.fields.size();
    std::vector<double> buf((size_t)N * size_one);
    // compute_theta / compute_usfe of the apps that have them (system 0 for
    // theta, system n-1 of material->A_mpea for usfeN); app.h:130-131: 0 otherwise
    static const char *theta_apps[] = {"1L1S111Bbcc", "1L1S100B", "1L1S110Binclined", "1SBCC", "3SBCC",
                                       "BCCEdge", "BCCEdgePlanar", "cond100B", "FrankRead", "frank_read"};
    static const char *usfe_apps[] = {"1SBCC", "3SBCC", "BCCEdge", "BCCEdgePlanar"};
    bool has_theta = false, has_usfe = false;
    for (auto a : theta_apps) has_theta |= app.style == a;
    for (auto a : usfe_apps) has_usfe |= app.style == a;
    std::vector<double> th;
    for (auto &f : d.fields)
        if (f == "theta" && has_theta && th.empty()) fft.theta(th);
    for (long long r = 0; r < N; ++r) {
        const int i = (int)(r / ((long long)fft.ny * fft.nz)), j = (int)((r / fft.nz) % fft.ny), k = (int)(r % fft.nz);
        for (int c = 0; c < size_one; ++c) {
            const std::string &f = d.fields[c];
            double v = 0.0;
            if (f == "x") v = i;
            else if (f == "y") v = j;
            else if (f == "z") v = k;
            else if (f.rfind("xir", 0) == 0 || f.rfind("xii", 0) == 0) {
                const int a = std::atoi(f.c_str() + 3) - 1;
                v = fft.xi[2 * (r + (long long)a * N) + (f[2] == 'r' ? 0 : 1)];
            } else if (f == "theta") {
                v = th.empty() ? 0.0 : th[r];
            } else if (f.rfind("usfe", 0) == 0) {
                const long long q = r + (long long)(f[4] - '1') * N;
                v = (has_usfe && q < (long long)fft.mat.A_mpea.size()) ? fft.mat.A_mpea[q] : 0.0;
            } else if (f == "deltao" || f == "ddelta") {
                // App1L1S100B / Appcond100B::compute_delta, compute_ddelta
                // (app_1L1S_100B.cpp:348-389); other apps return 0 (app.h:132-133)
                const std::vector<double> &src = f == "deltao" ? fft.delta : fft.ddelta;
                const bool has = app.style == "1L1S100B" || app.style == "cond100B";
                v = (has && !src.empty()) ? src[r] : 0.0;
            } else if (f.size() == 3 && (f[0] == 'p' || f[0] == 'e')) {
                // AppFCC::compute_stress / compute_strain (app_fcc.cpp:438-502): component
                // (i, j) of the last computed data_sigma / data_eps.  As in the reference,
                // dumps are written before the diagnostics of the same output step.
                const int ii = f[1] - 'x', jj = f[2] - 'x';
                const std::vector<double> &src = f[0] == 'p' ? fft.data_sigma : fft.data_eps;
                v = src.empty() ? 0.0 : src[2 * (r + (long long)ii * N + (long long)jj * 3 * N)];
            }
            buf[(size_t)r * size_one + c] = v;
        }
    }
    if (d.binary) {
        const int ndump = (int)N, nprocs = 1;
        const double box[6] = {0.0, (double)fft.nx, 0.0, (double)fft.ny, 0.0, (double)fft.nz};
        fwrite(&d.idump, sizeof(int), 1, fp);
        fwrite(&time, sizeof(double), 1, fp);
        fwrite(&ndump, sizeof(int), 1, fp);
        fwrite(box, sizeof(double), 6, fp);
        fwrite(&size_one, sizeof(int), 1, fp);
        fwrite(&nprocs, sizeof(int), 1, fp);
        const int n = ndump * size_one;
        fwrite(&n, sizeof(int), 1, fp);
        fwrite(buf.data(), sizeof(double), buf.size(), fp);
    } else {
        std::string cols;
        for (auto &f : d.fields) cols += f + " ";
        if (!cols.empty()) cols.pop_back();
        fprintf(fp, "ITEM: TIMESTEP\n%d %10g\nITEM: NUMBER OF ATOMS\n%lld\nITEM: BOX BOUNDS\n", d.idump, time, N);
        fprintf(fp, "%g %g\n%g %g\n%g %g\n", 0.0, app.scale[0] * fft.nx, 0.0, app.scale[1] * fft.ny, 0.0,
                app.scale[2] * fft.nz);
        fprintf(fp, "ITEM: ATOMS %s\n", cols.c_str());
        // DumpText::write_text (:385-400): "%g " per value, then a newline
        std::string line;
        char num[64];
        for (long long r = 0; r < N; ++r) {
            line.clear();
            for (int c = 0; c < size_one; ++c) {
                snprintf(num, sizeof num, "%g ", buf[(size_t)r * size_one + c]);
                line += num;
            }
            fprintf(fp, "%s\n", line.c_str());
        }
    }
    ++d.idump;
    fclose(fp);
}

double Output::setup(App &app, FFT_HIP &fft, double time)
{
    // Output::setup (output.cpp:79-126): first dumps, stats-driven diagnostics
    // on the initial field, stats header and first line.
    double dump_time = app.stoptime;
    for (auto &d : dumps) {
        write_dump(d, app, fft, time);
        d.next_time = next_time(time, d.every);
        dump_time = std::min(dump_time, d.next_time);
    }
    if (diag_strain || diag_stress) fft.diagnostics();
    stats_header();
    stats(app, fft, 0);
    stats_time = app.stoptime;
    if (stats_delta > 0.0) stats_time = next_time(time, stats_delta);
    return std::min({app.stoptime, dump_time, stats_time});
}

double Output::compute(App &app, FFT_HIP &fft, double time, int done)
{
    // Output::compute (output.cpp:133-188)
    double dump_time = app.stoptime;
    for (auto &d : dumps) {
        if (time >= d.next_time || done) {
            write_dump(d, app, fft, time);
            d.next_time = next_time(time, d.every);
        }
        dump_time = std::min(dump_time, d.next_time);
    }
    const bool sflag = time >= stats_time || done;
    if (sflag) {
        if (diag_strain || diag_stress) fft.diagnostics();
        stats(app, fft, 1);
        stats_time = app.stoptime;
        if (stats_delta > 0.0) stats_time = next_time(time, stats_delta);
    }
    return std::min({app.stoptime, dump_time, stats_time});
}

// ========================================================================= App
void App::init_app(FFT_HIP &fft)
{
    // init_app of the app style: set_slip (+ rotation), initial_sxtal, stopsteps
    if (!slip_table(style, fft.slip)) deck_error("Unrecognized app style " + style);
    rotate_slip(fft.slip, fft.geo);
    fft.slip_systems = fft.slip.slip_systems;
    fft.num_planes = fft.slip.num_planes;
    fft.CD = CD >= 0 ? CD : fft.slip.CD;
    fft.timestep = timestep >= 0 ? timestep : fft.slip.timestep;
    fft.core_flag = core_flag >= 0 ? core_flag : fft.slip.default_core;
    if (fft.core_flag == 1 || fft.core_flag == 2 || fft.core_flag == 3 || fft.core_flag == 4 || fft.core_flag == 5 ||
        fft.core_flag == 9) {
        fft.coef = core_coefficients(fft.mat, fft.core_flag);
        fft.ncoef = (int)fft.coef.size();
    }
    if (fft.core_flag == 6) {
        double sr[3][3];
        pfdd_host::rotate_stress(fft.geo, fft.sigma, sr);
        fft.coef = {usfe_angle_tau(fft.mat, sr, fft.angle_to_110)};
        fft.ncoef = 1;
    }
    if (fft.core_flag == 2 && fft.slip_systems % 3)
        deck_error("core_energy 2 couples the slip systems of each plane in threes");
    if (fft.core_flag == 7 && fft.mat.A_mpea.empty()) deck_error("core_energy 7 needs 'material ... mpea FILE'");
    if (fft.core_flag == 7) {
        // Material::read_mpea_file (material.cpp:848-853): the header must match the box
        const int *d = fft.mat.mpea_dims;
        if (d[0] != fft.nx || d[1] != fft.ny || d[2] != fft.nz || d[3] != fft.slip_systems) {
            char buf[256];
            snprintf(buf, sizeof(buf), "MPEA core energy file size does not match box %d %d %d %d %d %d %d %d",
                     fft.nx, d[0], fft.ny, d[1], fft.nz, d[2], fft.slip_systems, d[3]);
            deck_error(buf);
        }
    }
    Grid g{fft.nx, fft.ny, fft.nz, fft.slip_systems};
    const long long n = (long long)g.S * g.N();
    fft.xi.assign(2 * n, 0.0);
    fft.xo.assign(n, 0.0);
    // the app's default setup_flag when the deck gives no initial command
    if ((!init_set || fft.slip.force_init) && !fft.slip.default_init.empty()) init_style = fft.slip.default_init;
    if (!init_set && fft.slip.default_init.empty() && !fft.slip.missing_init.empty())
        printf("WARNING: the default setup %s of app %s is not built here; the initial field is zero "
               "(give an initial command or app_style ... setup)\n",
               fft.slip.missing_init.c_str(), style.c_str());
    if (init_style == "none") {
        // initial_sxtal_NoLoop (fft_fftw_slab.cpp:2196-2240): all zero
    } else if (init_sxtal_more(init_style, g, fft.geo, scale[0], (double)fft.nx, fft.xi.data(), fft.xo.data())) {
        // mode1, mode2, 1L*FCC, 2LOrtho, 3D3Shcp*
    } else if (init_style == "StraightOrtho") {
        init_sxtal_straight_ortho(g, fft.xi.data());
    } else if (init_style == "StraightNonOrtho") {
        init_sxtal_straight_nonortho(g, fft.xi.data());
    } else if (init_style == "1LOrthoBCC") {
        init_sxtal_1lorthobcc(g, scale[0], scale[1], fft.xi.data());
    } else if (init_style == "loops") {
        init_random_loops(g, fft.slip, init_nloops, init_rmin, init_rmax, init_noise, seed, fft.xi.data());
    } else if (init_style == "loop") {
        init_sxtal_loop(g, fft.xi.data());
    } else if (init_style == "frank_read") {
        init_sxtal_frank_read(g, fft.xi.data(), fft.xo.data());
    } else if (init_style == "edge") {
        init_sxtal_edge(g, fft.xi.data());
    } else if (init_style == "3SBCC") {
        init_sxtal_3sbcc(g, fft.xi.data());
    } else {
        deck_error("Illegal setup command: initial " + init_style);
    }
    for (const Disc &dsc : discs) {
        if (dsc.a < 0 || dsc.a >= g.S) deck_error("initial disc: slip system out of range");
        for (int i = 0; i < g.nx; ++i)
            for (int j = 0; j < g.ny; ++j)
                for (int k = 0; k < g.nz; ++k) {
                    const int p[3] = {i, j, k};
                    if (p[dsc.axis] != dsc.plane) continue;
                    const int u = dsc.axis == 0 ? j : i, v = dsc.axis == 2 ? j : k;
                    if ((u - dsc.c0) * (u - dsc.c0) + (v - dsc.c1) * (v - dsc.c1) > dsc.r * dsc.r) continue;
                    const long long idx = k + (long long)j * g.nz + (long long)i * g.ny * g.nz + (long long)dsc.a * g.N();
                    fft.xi[2 * idx] += 1.0;
                }
    }
    stopsteps = (int)std::floor(stoptime / fft.timestep);
}

// ===================================================================== SolveGL
void SolveGL::iterate(App &app, FFT_HIP &fft, Output &out)
{
    // SolveGL::iterate (solve_GL.cpp:70-158).  Consecutive iterations that
    // need no output are run as one GPU call; the device stops the call after
    // the first iteration with xinorm < tol, so the break semantics match.
    const double t0 = wall();
    printf("Beginning Time Evolution\n");
    // init_loop zeroes delta and ddelta at the start of every iteration and
    // project_core_energy (non_Schmid runs) refills them after the output of
    // the iteration (solve_GL.cpp:87,119-122): dumps taken during the run see
    // zeros; the values of the last iteration reach the setup dump of a
    // following run command
    bool any = false;
    for (int sstate = 0; sstate < app.stopsteps; ++sstate) {
        fft.rotate_stress();
        fft.resolSS();
        if (!any) {
            fft.delta.clear();
            fft.ddelta.clear();
            any = true;
        }
        int it = 0;
        while (it < max_iter) {
            const double need = std::ceil(app.nextoutput - app.time); // first iteration that must output
            int k = (int)std::min<double>(std::max(0.0, need), (double)(max_iter - 1 - it)) + 1;
            k = std::max(1, k);
            const int done = fft.iterate_block(k, tol);
            ++gpu_calls;
            it += done;
            const int it_last = it - 1;
            app.time += done - 1; // time of the last executed iteration
            out.loop_time = wall() - t0;
            const bool conv = fft.xinorm < tol;
            if (conv || fft.ximin > 1 || it_last == max_iter - 1) {
                printf("printing output it %d\n", it_last);
                app.nextoutput = out.compute(app, fft, app.time, 1);
            } else if (app.time >= app.nextoutput) {
                app.nextoutput = out.compute(app, fft, app.time, 0);
            }
            app.time += 1;
            app.steps += done;
            if (conv) {
                printf("Loading step %d converged after %d iterations\n", sstate, it_last);
                break;
            }
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) fft.sigma[i][j] = fft.sigma[i][j] + fft.deltasig[i][j];
    }
    if (any && fft.non_schmid) fft.project_core_energy();
    out.loop_time = wall() - t0;
}

void SolveGL::iterate_cond(App &app, FFT_HIP &fft, Output &out)
{
    // SolveGLCond::iterate (solve_GL+cond.cpp:70-150) as it executes: the
    // body of the iteration loop ends in an unconditional break (the if of
    // :134 guards only prep_backward), so every stress state runs one
    // iteration -- the dislocation step and the conductivity step -- and
    // then prints backwardcond's conductivity once more, computed by
    // backwardcond on its own output.
    const double t0 = wall();
    printf("Beginning Time Evolution\n");
    for (int sstate = 0; sstate < app.stopsteps; ++sstate) {
        fft.rotate_stress();
        fft.resolSS();
        fft.delta.clear(); // init_loop
        fft.ddelta.clear();
        for (int it = 0; it < max_iter; ++it) {
            fft.cond_iterate();
            printf("\nconductivity_local: %f \n", fft.cond_last.conductivity);
            printf("\nconductivity: %f \n", fft.cond_last.conductivity);
            fft.iterate_block(1, 0.0);
            ++gpu_calls;
            out.loop_time = wall() - t0;
            if (app.time >= app.nextoutput) app.nextoutput = out.compute(app, fft, app.time, 0);
            app.time += 1;
            app.steps += 1;
            if (fft.non_schmid) fft.project_core_energy();
            printf("\nconductivity_local: %f \n", fft.cond_last.conductivity_repeat);
            printf("\nconductivity: %f \n", fft.cond_last.conductivity_repeat);
            break;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) fft.sigma[i][j] = fft.sigma[i][j] + fft.deltasig[i][j];
    }
    out.loop_time = wall() - t0;
}

// ======================================================================= Input
void Input::file(FILE *fp)
{
    // Input::file (input.cpp:117-190): '&' continues a line, '#' starts a comment
    std::string acc;
    char buf[4096];
    while (fgets(buf, sizeof buf, fp)) {
        std::string s = buf;
        while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
        size_t amp = s.find_last_not_of(" \t");
        if (amp != std::string::npos && s[amp] == '&') {
            acc += s.substr(0, amp) + " ";
            continue;
        }
        acc += s;
        line(acc);
        acc.clear();
    }
    if (!acc.empty()) line(acc);
}

void Input::line(const std::string &text)
{
    std::string s = text;
    const size_t hash = s.find('#');
    if (hash != std::string::npos) s = s.substr(0, hash);
    std::istringstream is(s);
    std::vector<std::string> w;
    std::string t;
    while (is >> t) w.push_back(t);
    if (!w.empty()) command(w);
}

void Input::command(const std::vector<std::string> &w)
{
    const std::string &c = w[0];
    if (c == "seed") {
        app.seed = (int)num(w, 1, "seed");
    } else if (c == "app_style") {
        if (w.size() < 2) deck_error("Illegal app_style command");
        app.style = w[1];
        SlipTable probe;
        if (!slip_table(app.style, probe)) deck_error("Unrecognized app style " + app.style);
        for (size_t i = 2; i < w.size(); ++i) {
            if (w[i] == "core_energy") {
                app.core_flag = (int)num(w, ++i, "app_style");
            } else if (w[i] == "setup" && i + 1 < w.size()) {
                // App::App / set_initial_sxtal (app.cpp:84-88, 810-846): the
                // setups built here; "screw" is initial_sxtal_3SBCC
                const std::string st = w[++i];
                static const char *built[] = {"loop",          "none",          "edge",          "screw",
                                              "mode1",         "mode2",         "1LOrthoFCC",    "1LNonOrthoFCC",
                                              "1LInclinedFCC", "1LOrthoBCC",    "2LOrtho",       "StraightOrtho",
                                              "StraightNonOrtho", "frank_read", "3D3ShcpNotch",  "3D3ShcpBasal"};
                bool ok = false;
                for (auto b : built) ok |= st == b;
                if (!ok) deck_error("setup " + st + " is not built in this driver (use the initial command)");
                app.init_style = st == "screw" ? "3SBCC" : st;
                app.init_set = true;
            } else {
                deck_error("Illegal app_style command: " + w[i]);
            }
        }
    } else if (c == "fft_style") {
        // FFT::FFT argument parsing (fft.cpp:96-153); style hip (or the
        // reference's fftw_slab, accepted and run on the GPU)
        if (w.size() < 2 || (w[1] != "hip" && w[1] != "fftw_slab")) deck_error("Invalid fft style");
        for (size_t i = 2; i < w.size(); ++i) {
            if (w[i] == "pbc") {
                for (int q = 0; q < 3; ++q)
                    if (num(w, ++i, "fft") != 1) deck_error("only periodic boxes are supported");
            } else if (w[i] == "mode") {
                fft.mode = (int)num(w, ++i, "fft");
            } else if (w[i] == "primitive") {
                double v[3][3];
                for (int a = 0; a < 3; ++a) {
                    if (++i >= w.size()) deck_error("Illegal fft command");
                    const char ax = "xyz"[a];
                    if (w[i] != std::string(1, ax)) deck_error("Illegal fft command: primitive");
                    for (int q = 0; q < 3; ++q) v[a][q] = num(w, ++i, "fft");
                }
                fft.geo.set_primitive(v[0], v[1], v[2]);
            } else if (w[i] == "precision") {
                ++i;
                if (i >= w.size()) deck_error("Illegal fft command");
                if (w[i] == "fp64") fft.precision = PFDD_FP64;
                else if (w[i] == "fp32") fft.precision = PFDD_FP32;
                else deck_error("precision must be fp64 or fp32");
            } else if (w[i] == "stored_B") {
                fft.flags |= PFDD_GL_FLAG_STORED_B;
            } else {
                deck_error("Illegal fft_style command: " + w[i]);
            }
        }
    } else if (c == "dimension") {
        if ((int)num(w, 1, "dimension") != 3) deck_error("only dimension 3 is supported");
    } else if (c == "lattice") {
        // lattice style is irrelevant for the structured periodic FFT grid
    } else if (c == "region") {
        if (w.size() < 9 || w[2] != "block") deck_error("Illegal region command");
        fft.nx = (int)(num(w, 4, "region") - num(w, 3, "region"));
        fft.ny = (int)(num(w, 6, "region") - num(w, 5, "region"));
        fft.nz = (int)(num(w, 8, "region") - num(w, 7, "region"));
    } else if (c == "create_box") {
        if (fft.nx <= 0) deck_error("Create_box region ID does not exist");
        box = true;
    } else if (c == "create_sites") {
        if (!box) deck_error("Cannot create sites before simulation box is defined");
        sites = true;
    } else if (c == "material") {
        // Material::Material (material.cpp:32-167).  FFT::set_material
        // (fft.cpp:221-246): the first material command is fft->material
        // (materialA), the second materialB (the conductivity layers); later
        // ones are parsed and kept by nobody on this path
        Material spare;
        Material &m = !have_material ? fft.mat : !fft.have_matB ? fft.matB : spare;
        const bool second = have_material && !fft.have_matB;
        for (size_t i = 1; i < w.size(); ++i) {
            const std::string &k = w[i];
            if (k == "preset") {
                if (i + 1 >= w.size() || !m.preset(w[++i])) deck_error("Unknown material preset");
            } else if (k == "mpea") {
                if (i + 1 >= w.size()) deck_error("Illegal material command");
                FILE *f = fopen(w[++i].c_str(), "r");
                if (!f) deck_error("Cannot open input file " + w[i]);
                int n1, n2, n3, ns;
                if (fscanf(f, "%d %d %d %d", &n1, &n2, &n3, &ns) != 4) deck_error("bad MPEA file header");
                m.A_mpea.assign((size_t)n1 * n2 * n3 * ns, 0.0);
                m.mpea_dims[0] = n1;
                m.mpea_dims[1] = n2;
                m.mpea_dims[2] = n3;
                m.mpea_dims[3] = ns;
                int a, b, cc, s;
                double v;
                while (fscanf(f, "%d %d %d %d %lf", &a, &b, &cc, &s, &v) == 5)
                    m.A_mpea[(size_t)a * n2 * n3 + (size_t)b * n3 + cc + (size_t)s * n1 * n2 * n3] = v * 1e-3;
                fclose(f);
            } else if (k == "El") {
                for (int q = 0; q < 3; ++q) m.El[q] = num(w, ++i, "material");
            } else {
                const double v = num(w, ++i, "material");
                if (k == "a") m.a = v;
                else if (k == "dlayer") m.dlayer = v;
                else if (k == "Nlayers") m.Nlayers = (int)v;
                else if (k == "Ngrains") m.Ngrains = (int)v;
                else if (k == "Rgrain") m.Rgrain = v;
                else if (k == "alpha_rho") m.alpha_resistivity = v, m.alpha_mfpel = v;
                else if (k == "pinterface") m.pinterface = v;
                else if (k == "rho_disloc") m.rho_disloc = v;
                else if (k == "beta_disloc") m.beta_disloc = v;
                else if (k == "b") m.b = v;
                else if (k == "mu") m.mu = v;
                else if (k == "young") m.young = v;
                else if (k == "isf") m.isf = v;
                else if (k == "usf") m.usf = v;
                else if (k == "An") m.An = v;
                else if (k == "Cn") m.Cn = v;
                else if (k == "c0") m.c0 = v;
                else if (k == "c1") m.c1 = v;
                else if (k == "c2") m.c2 = v;
                else if (k == "c3") m.c3 = v;
                else if (k == "c4") m.c4 = v;
                else if (k == "a1") m.a1 = v;
                else if (k == "a3") m.a3 = v;
                else if (k.size() == 3 && (k.compare(0, 2, "aa") == 0 || k.compare(0, 2, "bb") == 0) &&
                         k[2] >= '0' && k[2] <= '4' && !(k[0] == 'b' && k[2] == '0'))
                    m.pyr[k[0] == 'a' ? k[2] - '0' : 4 + (k[2] - '0')] = v;
                else deck_error("Illegal material command: " + k);
            }
        }
        if (m.b == 0.0 && m.a > 0.0) m.b = m.a / std::sqrt(2.0);
        if (!(m.mu > 0) || !(m.young > 0)) deck_error("material needs mu and young (or a preset)");
        m.derive();
        if (second) fft.have_matB = true;
        have_material = true;
    } else if (c == "solve_style") {
        if (w.size() < 2 || (w[1] != "GL" && w[1] != "GLCond")) deck_error("Invalid solve style");
        solve.cond = fft.cond = w[1] == "GLCond";
        for (size_t i = 2; i < w.size(); ++i) {
            if (w[i] == "max_iter") solve.max_iter = (int)num(w, ++i, "solve");
            else if (w[i] == "tol") solve.tol = num(w, ++i, "solve");
            else deck_error("Illegal solve command");
        }
    } else if (c == "diag_style") {
        if (w.size() < 2) deck_error("Illegal diag_style command");
        if (w[1] == "strain") output.diag_strain = true;
        else if (w[1] == "stress") output.diag_stress = true;
        else deck_error("Invalid diag_style " + w[1]);
    } else if (c == "temperature") {
        fft.temperature = num(w, 1, "temperature"); // app->temperature
    } else if (c == "non_Schmid") {
        // App::set_nonSchmid (app.cpp:446-460): angle_to_110 w1 w2 w3
        if (w.size() != 5) deck_error("Illegal non_Schmid command");
        fft.non_schmid = true;
        fft.angle_to_110 = num(w, 1, "non_Schmid");
        fft.w1 = num(w, 2, "non_Schmid");
        fft.w2 = num(w, 3, "non_Schmid");
        fft.w3 = num(w, 4, "non_Schmid");
    } else if (c == "scale") {
        // App::set_scale (app.cpp:383-444)
        for (int q = 0; q < 3; ++q) app.scale[q] = num(w, 1 + q, "scale");
        app.scale_set = true;
    } else if (c == "sigma") {
        // App::set_sigma (app.cpp:325-381): xx yy zz xy xz yz, units of mu
        if (w.size() <= 7) deck_error("Illegal sigma command");
        for (size_t i = 1; i < w.size();) {
            double(*t)[3] = w[i] == "initial" ? app.sigma0 : w[i] == "delta" ? app.dsigma : nullptr;
            if (!t) deck_error("Illegal sigma command");
            double v[6];
            for (int q = 0; q < 6; ++q) v[q] = num(w, i + 1 + q, "sigma");
            t[0][0] = v[0];
            t[1][1] = v[1];
            t[2][2] = v[2];
            t[0][1] = t[1][0] = v[3];
            t[0][2] = t[2][0] = v[4];
            t[1][2] = t[2][1] = v[5];
            i += 7;
        }
    } else if (c == "stats") {
        output.stats_delta = num(w, 1, "stats");
    } else if (c == "dump") {
        if (w.size() < 5 || w[2] != "text") deck_error("Illegal dump command");
        DumpSpec d;
        d.id = w[1];
        d.every = num(w, 3, "dump");
        d.file = w[4];
        d.binary = d.file.size() > 4 && d.file.compare(d.file.size() - 4, 4, ".bin") == 0;
        for (size_t i = 5; i < w.size(); ++i) {
            const std::string &f = w[i];
            static const char *tens[12] = {"pxx", "pyy", "pzz", "pxy", "pxz", "pyz",
                                           "exx", "eyy", "ezz", "exy", "exz", "eyz"};
            bool t = false;
            for (auto q : tens) t |= f == q;
            const bool ok = t || f == "x" || f == "y" || f == "z" || f.rfind("xir", 0) == 0 ||
                            f.rfind("xii", 0) == 0 || f == "deltao" || f == "ddelta" || f == "theta" ||
                            f == "usfe1" || f == "usfe2" || f == "usfe3";
            if (!ok)
                deck_error("dump field " + f +
                           " is not available (x y z theta usfe1-3 deltao ddelta pxx..pyz exx..eyz xirN xiiN)");
            if (t) fft.want_fields = true;
            d.fields.push_back(f);
        }
        output.dumps.push_back(d);
    } else if (c == "run") {
        run(w);
    } else if (c == "initial") {
        // extension: seeded initial fields
        if (w.size() < 2) deck_error("Illegal initial command");
        app.init_set = true;
        if (w[1] == "loops") {
            app.init_style = "loops";
            app.init_nloops = (int)num(w, 2, "initial");
            app.init_rmin = num(w, 3, "initial");
            app.init_rmax = num(w, 4, "initial");
            app.init_noise = num(w, 5, "initial");
        } else if (w[1] == "noise") {
            app.init_style = "loops";
            app.init_nloops = 0;
            app.init_noise = num(w, 2, "initial");
        } else if (w[1] == "disc") {
            App::Disc d{(int)num(w, 2, "initial"), (int)num(w, 3, "initial"), (int)num(w, 4, "initial"),
                        num(w, 5, "initial"), num(w, 6, "initial"), num(w, 7, "initial")};
            if (d.axis < 0 || d.axis > 2) deck_error("initial disc: axis must be 0, 1 or 2");
            app.discs.push_back(d);
        } else if (w[1] == "sxtal") {
            if (w.size() < 3) deck_error("Illegal initial command");
            app.init_style = w[2];
        } else {
            deck_error("Illegal initial command");
        }
    } else if (c == "mobility") {
        // extension: CD and timestep (hard-coded per app in the reference)
        app.CD = num(w, 1, "mobility");
        app.timestep = num(w, 2, "mobility");
    } else if (c == "write_xi") {
        if (w.size() < 2) deck_error("Illegal write_xi command");
        write_xi_file = w[1];
    } else {
        deck_error("Unknown command: " + c);
    }
}

void Input::print_setup()
{
    printf("grid %d %d %d slip_systems %d core_flag %d CD %.17g timestep %.17g\n", fft.nx, fft.ny, fft.nz,
           fft.slip_systems, fft.core_flag, fft.CD, fft.timestep);
    printf("four");
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) printf(" %.17g", fft.geo.four[r][c]);
    printf("\nmaterial C44 %.17g C12 %.17g C11 %.17g S12 %.17g S44 %.17g mu_s %.17g b %.17g\n", fft.mat.C44,
           fft.mat.C12, fft.mat.C11, fft.mat.S12, fft.mat.S44, fft.mat.mu_s, fft.mat.b);
    printf("coef");
    for (double v : fft.coef) printf(" %.17g", v);
    printf("\n");
    for (int a = 0; a < fft.slip_systems; ++a)
        printf("slip %d xn %.17g %.17g %.17g xb %.17g %.17g %.17g\n", a, fft.slip.xn[3 * a], fft.slip.xn[3 * a + 1],
               fft.slip.xn[3 * a + 2], fft.slip.xb[3 * a], fft.slip.xb[3 * a + 1], fft.slip.xb[3 * a + 2]);
    rotate_stress(fft.geo, fft.sigma, fft.sigma_rot);
    std::vector<double> tau(fft.slip_systems);
    if (fft.non_schmid)
        resolss_non_schmid(fft.slip, fft.sigma_rot, fft.mat.dslip, fft.angle_to_110, fft.w1, fft.w2, fft.w3, tau.data());
    else
        resolss(fft.slip, fft.sigma_rot, fft.mat.dslip, tau.data());
    printf("tau");
    for (double v : tau) printf(" %.17g", v);
    printf("\nstopsteps %d max_iter %d tol %.17g\n", app.stopsteps, solve.max_iter, solve.tol);
}

void Input::run(const std::vector<std::string> &w)
{
    // App::run (app.cpp:127-205): init, setup, iterate, finish
    if (!box || !sites) deck_error("Cannot run application until simulation box is defined");
    if (!have_material) deck_error("material must be defined before run");
    if (app.style.empty()) deck_error("app_style must be defined before run");
    if (!app.scale_set) deck_error("scale must be set before run (it defines the reciprocal lattice)");
    app.stoptime = app.time + num(w, 1, "run");
    if (app.stoptime < app.time) deck_error("Illegal run command");
    if (run_count++ == 0) {
        fft.geo.set_scale(app.scale[0], app.scale[1], app.scale[2]);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                fft.sigma[i][j] = app.sigma0[i][j];
                fft.deltasig[i][j] = app.dsigma[i][j];
            }
        app.init_app(fft);
        for (auto &d : output.dumps)
            for (auto &f : d.fields)
                if (f.rfind("xir", 0) == 0 || f.rfind("xii", 0) == 0) {
                    const int a = std::atoi(f.c_str() + 3);
                    if (a < 1 || a > fft.slip_systems) deck_error("dump field " + f + ": no such slip system");
                }
        if (check_only) {
            print_setup();
            return;
        }
        fft.init();
        if (fft.cond) fft.cond_init();
    } else {
        app.stopsteps = (int)std::floor(app.stoptime / fft.timestep);
        if (check_only) return;
    }
    printf("Setting up run ...\n");
    fft.rotate_stress(); // FFTW_Slab::setup (fft_fftw_slab.cpp:111-119)
    app.nextoutput = output.setup(app, fft, app.time);
    if (solve.cond) solve.iterate_cond(app, fft, output);
    else solve.iterate(app, fft, output);
    printf("Loop time of %g on %d procs\n", output.loop_time, 1);
    if (!write_xi_file.empty()) {
        fft.sync_host();
        FILE *f = fopen(write_xi_file.c_str(), "wb");
        if (!f) deck_error("Cannot open " + write_xi_file);
        fwrite(fft.xi.data(), sizeof(double), fft.xi.size(), f);
        fclose(f);
    }
}

} // namespace pfdd_host
