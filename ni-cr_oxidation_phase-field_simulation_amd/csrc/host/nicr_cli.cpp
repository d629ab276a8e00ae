// nicr_cli.cpp -- drop-in programs nicr2d / nicr3d (NICR_CLI_DIM = 2 or 3).
//
// Same command line, output directory policy, output files and exit codes as the
// reference programs (run(), mpi2d/NiCr2D_grains.cpp:1161-1246 and
// mpi3d-slab/NiCr3D_grains.cpp:1309-1395). The time step runs on the GPU through
// libnicr (include/nicr.h); the host batches the steps between events (Ksp
// scans in physical time, output cadence, progress lines) into one nicr_step
// call, so it synchronises with the device only at those events.
#include "nicr_host.hpp"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <filesystem>
#include <functional>
#include <thread>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#ifndef NICR_CLI_DIM
#define NICR_CLI_DIM 3
#endif

namespace fs = std::filesystem;
using namespace nicrh;

namespace {

void check(int rc, const char* what) {
    if (rc != NICR_OK) throw std::runtime_error(std::string(what) + ": " + nicr_last_error());
}

struct Device {
    void* ctx = nullptr;
    ~Device() {
        if (ctx) nicr_destroy(ctx);
    }
};

std::string backend_name() {
    hipDeviceProp_t prop{};
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        return std::string("hip_") + prop.gcnArchName + "_single_gpu";
    return "hip_single_gpu";
}

// Wall time per phase of the run, summed over the run and printed to stderr at
// exit when the environment sets NICR_PHASE_TIMES; the outputs do not change.
struct PhaseTimes {
    bool on = std::getenv("NICR_PHASE_TIMES") != nullptr;
    std::vector<std::pair<std::string, double>> acc;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char* name) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        add(name, std::chrono::duration<double>(now - t).count());
        t = now;
    }
    void add(const std::string& name, double d) {
        if (!on) return;
        for (auto& a : acc)
            if (a.first == name) { a.second += d; return; }
        acc.emplace_back(name, d);
    }
    ~PhaseTimes() {
        for (const auto& a : acc) std::fprintf(stderr, "phase %-10s %9.3f s\n", a.first.c_str(), a.second);
    }
};

// One output snapshot (diagnostics row, grain rows, visualisation file) written
// on a host thread while the device runs the next steps. The host arrays of
// the snapshot stay untouched until wait() joins the thread, which the run
// loop does before the next download, so rows reach the files in step order
// and every file holds what the synchronous writer would have written. A
// write error surfaces at the next wait().
struct OutputThread {
    std::thread th;
    std::exception_ptr err;
    std::vector<std::pair<std::string, double>> times;   // writer phases of the last snapshot
    void start(std::function<void(OutputThread&)> fn) {
        th = std::thread([this, fn] {
            try {
                fn(*this);
            } catch (...) {
                err = std::current_exception();
            }
        });
    }
    void timed(const char* name, const std::function<void()>& fn) {
        const auto t0 = std::chrono::steady_clock::now();
        fn();
        times.emplace_back(name, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    void wait(PhaseTimes& pt) {
        if (th.joinable()) th.join();
        for (const auto& t : times) pt.add(t.first, t.second);
        times.clear();
        if (err) {
            std::exception_ptr e = err;
            err = nullptr;
            std::rethrow_exception(e);
        }
    }
    void wait_quietly() {
        if (th.joinable()) th.join();
        err = nullptr;
        times.clear();
    }
    ~OutputThread() {
        if (th.joinable()) th.join();
    }
};

long double join(const double v[2]) { return (long double)v[0] + (long double)v[1]; }

void upload(const HostModel& m, void* ctx) {
    check(nicr_upload_field(ctx, NICR_FIELD_O, m.o.data()), "upload o");
    check(nicr_upload_field(ctx, NICR_FIELD_CR, m.cr.data()), "upload cr");
    check(nicr_upload_field(ctx, NICR_FIELD_PHI, m.p.data()), "upload phi");
    check(nicr_upload_field(ctx, NICR_FIELD_GB, m.gb.data()), "upload gb");
    for (std::size_t g = 0; g < m.eta.size(); ++g)
        check(nicr_upload_field(ctx, NICR_FIELD_ETA0 + int(g), m.eta[g].data()), "upload eta");
    nicr_budget b{};
    b.nucO[0] = double(m.b.nucO);
    b.nucO[1] = double(m.b.nucO - (long double)b.nucO[0]);
    b.nucCr[0] = double(m.b.nucCr);
    b.nucCr[1] = double(m.b.nucCr - (long double)b.nucCr[0]);
    b.nuclei = m.b.nuclei;
    check(nicr_set_budget(ctx, &b), "set budget");
}

// Model::refresh + the fields every writer reads (2D:600, SL:830).
void download(HostModel& m, void* ctx, bool grain_terms = true) {
    check(nicr_download_field(ctx, NICR_FIELD_O, m.o.data()), "download o");
    check(nicr_download_field(ctx, NICR_FIELD_CR, m.cr.data()), "download cr");
    check(nicr_download_field(ctx, NICR_FIELD_PHI, m.p.data()), "download phi");
    for (std::size_t g = 0; g < m.eta.size(); ++g)
        check(nicr_download_field(ctx, NICR_FIELD_ETA0 + int(g), m.eta[g].data()), "download eta");
    check(nicr_refresh(ctx), "refresh");
    m.muO.resize(m.ncell);
    m.muCr.resize(m.ncell);
    m.mo.resize(m.ncell);
    m.mc.resize(m.ncell);
    check(nicr_download_field(ctx, NICR_FIELD_MU_O, m.muO.data()), "download mu_O");
    check(nicr_download_field(ctx, NICR_FIELD_MU_CR, m.muCr.data()), "download mu_Cr");
    check(nicr_download_field(ctx, NICR_FIELD_M_O, m.mo.data()), "download M_O");
    check(nicr_download_field(ctx, NICR_FIELD_M_CR, m.mc.data()), "download M_Cr");
    check(nicr_download_field(ctx, NICR_FIELD_GB, m.gb.data()), "download gb");
    if (m.has_eta()) {
        m.eta_s2.resize(m.ncell);
        m.eta_pairs.resize(m.ncell);
        check(nicr_download_field(ctx, NICR_FIELD_ETA_S2, m.eta_s2.data()), "download eta_s2");
        check(nicr_download_field(ctx, NICR_FIELD_ETA_PAIRS, m.eta_pairs.data()), "download eta_pairs");
    }
    // the grain rows' per-cell terms of this snapshot, from the device
    // (NICR_GRAIN_TERMS=host keeps the host loops)
    m.grain_terms_valid = false;
    const char* gt = std::getenv("NICR_GRAIN_TERMS");
    if (grain_terms && m.has_eta() && m.dim == 3 && m.eta.size() <= 256 && !(gt && std::strcmp(gt, "host") == 0)) {
        m.grain_terms.resize(m.ncell);
        m.grain_counts.resize(m.eta.size() + 1);
        check(nicr_grain_terms(ctx, m.grain_terms.data(), m.grain_extrema, m.grain_counts.data()), "grain terms");
        m.grain_terms_valid = true;
    }
    nicr_budget b{};
    check(nicr_get_budget(ctx, &b), "budget");
    m.b.nucO = join(b.nucO);
    m.b.nucCr = join(b.nucCr);
    m.b.ckO = join(b.ckO);
    m.b.ckCr = join(b.ckCr);
    m.b.clipO = join(b.clipO);
    m.b.clipCr = join(b.clipCr);
    m.b.bcO = join(b.bcO);
    m.b.bcCr = join(b.bcCr);
    m.b.nuclei = b.nuclei;
    m.b.clipped = b.clipped;
}

std::string step_message(const nicr_step_report& r, int dim) {
    std::ostringstream msg;
    const std::string cell = std::to_string(r.i) + "," + std::to_string(r.j) + (dim == 3 ? "," + std::to_string(r.k) : "");
    if (r.status == NICR_STEP_NONFINITE) return "Nonfinite trial field: reduce dt / inspect the model; no update accepted.";
    if (r.status == NICR_STEP_OUT_OF_BOUNDS) {
        msg << std::scientific << std::setprecision(16) << "Out-of-bounds trial at (" << cell << "): O=" << r.o
            << ", Cr=" << r.cr << ", phi=" << r.phi
            << ". No clipping/update accepted. Reduce dt; if failure persists, investigate model.";
        return msg.str();
    }
    msg << std::setprecision(17) << "Eta trial outside [0,1]: grain=" << r.grain << " cell=(" << cell
        << ") value=" << r.eta_value << ". All PDE fields rejected together; eta is never clipped. Check eta coefficients/dt.";
    return msg.str();
}

// ---------------------------------------------------------------------------
// --self-test: closed-form and invariant checks of the device step
// ---------------------------------------------------------------------------
void require(bool v, const char* msg) {
    if (!v) throw std::runtime_error(std::string("Self-test failed: ") + msg);
}

struct Ctx {
    void* h = nullptr;
    explicit Ctx(nicr_config c) { check(nicr_create(&c, &h), "nicr_create"); }
    ~Ctx() { nicr_destroy(h); }
};

nicr_config small(int dim, const char* bc) {
    nicr_config c;
    nicr_config_default(&c, dim);
    c.nx = dim == 2 ? 31 : 23;
    c.ny = dim == 2 ? 25 : 13;
    c.nz = dim == 2 ? 1 : 11;
    c.bc = std::string(bc) == "mu" ? NICR_BC_MU : (std::string(bc) == "closed" ? NICR_BC_CLOSED : NICR_BC_CONCENTRATION);
    c.ck = NICR_CK_OFF;
    c.nucleation = NICR_NUC_OFF;
    c.dt = 1.e-12;
    c.grains = 0;
    nicr_config_derive(&c);
    return c;
}

void smooth(int dim, const nicr_config& c, std::vector<double>& o, std::vector<double>& cr, std::vector<double>& p) {
    const std::size_t n = std::size_t(c.nx) * c.ny * (dim == 3 ? c.nz : 1);
    o.resize(n);
    cr.resize(n);
    p.resize(n);
    std::size_t q = 0;
    for (int i = 1; i <= c.nx; ++i)
        for (int j = 1; j <= c.ny; ++j)
            for (int k = 1; k <= (dim == 3 ? c.nz : 1); ++k, ++q) {
                const double kk = dim == 3 ? k : 0.;
                o[q] = .001 * (1. + .08 * std::sin(.3 * i + .5 * j - .4 * kk));
                cr[q] = .19 + .001 * std::cos(.2 * i - .3 * j + .6 * kk);
                p[q] = .1 + .01 * std::cos(.1 * i + .4 * j - .2 * kk);
            }
}

long double dsum(void* h, int field) {
    nicr_stats s{};
    check(nicr_field_stats(h, field, &s), "stats");
    return (long double)s.sum[0] + s.sum[1];
}

void self_test(int dim) {
    std::vector<double> o, cr, p;
    // closed boundaries conserve O and Cr (2D:825-826, SL:1160-1161)
    {
        nicr_config c = small(dim, "closed");
        Ctx d(c);
        smooth(dim, c, o, cr, p);
        check(nicr_upload_field(d.h, NICR_FIELD_O, o.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_CR, cr.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_PHI, p.data()), "up");
        const long double so = dsum(d.h, NICR_FIELD_O), sc = dsum(d.h, NICR_FIELD_CR);
        nicr_step_report r{};
        check(nicr_step(d.h, 5, &r), "step");
        require(r.status == 0 && r.steps_done == 5, "closed step accepted");
        require(std::abs(dsum(d.h, NICR_FIELD_O) - so) < 1.e-12L * (1 + so), "closed transport conserves O");
        require(std::abs(dsum(d.h, NICR_FIELD_CR) - sc) < 1.e-12L * (1 + sc), "closed transport conserves Cr");
    }
    std::cout << "PASS: closed-boundary O and Cr conservation on the device.\n";
    // O diffusion off freezes O exactly on every BC (2D:853-869, SL:1165-1170)
    for (const char* bc : {"mu", "closed", "concentration"}) {
        nicr_config c = small(dim, bc);
        c.oxygen_diffusion = 0;
        Ctx d(c);
        smooth(dim, c, o, cr, p);
        check(nicr_upload_field(d.h, NICR_FIELD_O, o.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_CR, cr.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_PHI, p.data()), "up");
        nicr_step_report r{};
        check(nicr_step(d.h, 3, &r), "step");
        std::vector<double> o2(o.size());
        check(nicr_download_field(d.h, NICR_FIELD_O, o2.data()), "down");
        nicr_budget b{};
        check(nicr_get_budget(d.h, &b), "budget");
        require(o2 == o && b.bcO[0] == 0., "O off freezes O transport/inlet on every BC");
    }
    std::cout << "PASS: oxygen switch freezes O for mu, closed and concentration boundaries.\n";
    // hard seeds: 97-cell R=5.5 disk (2D:829), 739-voxel R=5.5 sphere (SL:1192-1193)
    {
        nicr_config c = small(dim, "closed");
        if (dim == 3) { c.nx = 31; c.ny = 25; c.nz = 23; }
        c.radius_grid = 5.5;
        Ctx d(c);
        int n = 0;
        check(dim == 2 ? nicr_stamp(d.h, 15, 12, 1, &n) : nicr_stamp(d.h, 16, 13, 12, &n), "stamp");
        require(n == (dim == 2 ? 97 : 739), dim == 2 ? "R=5.5 disk has 97 sites" : "radius 5.5 sphere contains 739 voxels");
    }
    std::cout << "PASS: hard-seed geometry (" << (dim == 2 ? "97-cell disk" : "739-voxel sphere") << ").\n";
    // bounds=stop rejects the whole step and leaves the state and budget (2D:990-994)
    {
        nicr_config c = small(dim, "concentration");
        c.bounds = NICR_BOUNDS_STOP;
        Ctx d(c);
        smooth(dim, c, o, cr, p);
        cr[cr.size() - 3] = -0.1;
        check(nicr_upload_field(d.h, NICR_FIELD_O, o.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_CR, cr.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_PHI, p.data()), "up");
        nicr_step_report r{};
        check(nicr_step(d.h, 2, &r), "step");
        require(r.status == NICR_STEP_OUT_OF_BOUNDS && r.steps_done == 0, "trial failure rejects the step");
        std::vector<double> cr2(cr.size());
        check(nicr_download_field(d.h, NICR_FIELD_CR, cr2.data()), "down");
        nicr_budget b{};
        check(nicr_get_budget(d.h, &b), "budget");
        require(cr2 == cr && b.bcO[0] == 0., "rejected update leaves state and budget unchanged");
    }
    std::cout << "PASS: all-or-nothing rejection of an out-of-bounds trial.\n";
    // mu inlet face inventory balance (SL:1171-1172)
    {
        nicr_config c = small(dim, "mu");
        Ctx d(c);
        std::vector<double> z(std::size_t(c.nx) * c.ny * c.nz, 0.), cr0(z.size(), .1986);
        check(nicr_upload_field(d.h, NICR_FIELD_O, z.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_CR, cr0.data()), "up");
        check(nicr_upload_field(d.h, NICR_FIELD_PHI, z.data()), "up");
        nicr_step_report r{};
        check(nicr_step(d.h, 1, &r), "step");
        nicr_budget b{};
        check(nicr_get_budget(d.h, &b), "budget");
        const long double so = dsum(d.h, NICR_FIELD_O), bc = (long double)b.bcO[0] + b.bcO[1];
        require(so > 0 && std::abs(so - bc) <= 1.e-11L * (1 + std::abs(bc)), "mu inlet face inventory balance");
    }
    std::cout << "PASS: mu-BC inlet inventory balance.\n";
}

// ---------------------------------------------------------------------------
// --self-test, evolving grains: the eta checks of the reference self-tests
// (2D:1011-1093, SL:1201-1230) on the device step
// ---------------------------------------------------------------------------
HostConfig eta_config(int dim) {
    HostConfig hc;
    hc.dim = dim;
    nicr_config_default(&hc.c, dim);
    nicr_config& c = hc.c;
    c.grain_evolution = 1;
    c.grains = 1;
    c.nx = dim == 2 ? 35 : 23;
    c.ny = dim == 2 ? 29 : 13;
    c.nz = dim == 2 ? 1 : 11;
    c.grain_diameter_grid = dim == 2 ? 16. : 12.;
    c.bc = NICR_BC_CLOSED;
    c.ck = NICR_CK_OFF;
    c.nucleation = NICR_NUC_OFF;
    c.bounds = NICR_BOUNDS_LEGACY;
    c.dt = 1.e-10;
    hc.bc = "closed";
    hc.ck = "off";
    hc.nucleation = "off";
    hc.bounds = "legacy";
    nicr_config_derive(&c);
    return hc;
}

// the device's mask of the host arrays' eta, then the host's discrete grain
// energy (2D:367-381, SL:546-565) of that state
long double device_energy(HostModel& m, void* h) {
    for (std::size_t g = 0; g < m.eta.size(); ++g)
        check(nicr_upload_field(h, NICR_FIELD_ETA0 + int(g), m.eta[g].data()), "upload eta");
    check(nicr_upload_field(h, NICR_FIELD_PHI, m.p.data()), "upload phi");
    m.eta_s2.resize(m.ncell);
    m.eta_pairs.resize(m.ncell);
    check(nicr_download_field(h, NICR_FIELD_ETA_S2, m.eta_s2.data()), "download S");
    check(nicr_download_field(h, NICR_FIELD_ETA_PAIRS, m.eta_pairs.data()), "download P");
    m.grain_terms_valid = false;
    return m.grain_energy();
}

void eta_self_test(int dim) {
    // (1) the host-built eta state partitions unity after the trip through the device (2D:1015-1018, SL:1202-1205)
    {
        HostConfig hc = eta_config(dim);
        HostModel m(hc);
        nicr_config dc = hc.c;
        dc.n_eta = int(m.eta.size());
        require(dc.n_eta > 1, "several eta fields");
        Ctx d(dc);
        upload(m, d.h);
        std::vector<std::vector<double>> e(m.eta.size(), std::vector<double>(m.ncell));
        for (std::size_t g = 0; g < e.size(); ++g)
            check(nicr_download_field(d.h, NICR_FIELD_ETA0 + int(g), e[g].data()), "download eta");
        for (std::size_t q = 0; q < m.ncell; ++q) {
            double sum = 0.;
            for (const auto& f : e) sum += f[q];
            require(std::abs(sum - 1.) < 1.e-14, "eta initialization partitions unity");
        }
        // (2) eta-L = 0 freezes eta (SL:1206-1210; 2D runs the trial, which leaves it unchanged)
        nicr_config fc = dc;
        fc.eta_L = 0.;
        Ctx f(fc);
        upload(m, f.h);
        nicr_step_report r{};
        check(nicr_step(f.h, 5, &r), "step");
        require(r.status == 0 && r.steps_done == 5, "frozen-eta steps accepted");
        std::vector<double> v(m.ncell);
        for (std::size_t g = 0; g < m.eta.size(); ++g) {
            check(nicr_download_field(f.h, NICR_FIELD_ETA0 + int(g), v.data()), "download eta");
            require(v == m.eta[g], "eta-L=0 freezes eta");
        }
        int n = 0;
        check(dim == 2 ? nicr_stamp(f.h, 12, 1, 1, &n) : nicr_stamp(f.h, 12, 1, 1, &n), "stamp");
        require(n > 0, "hard seed inserted");
        for (std::size_t g = 0; g < m.eta.size(); ++g) {
            check(nicr_download_field(f.h, NICR_FIELD_ETA0 + int(g), v.data()), "download eta");
            require(v[m.at(12, 1, 1)] == 0., "hard seeds consume eta even with eta-L=0");
        }
    }
    std::cout << "PASS: eta initialization partitions unity; eta-L=0 freezes eta; seeds consume eta.\n";
    // (3) the device's eta force is the variational derivative of the discrete
    // grain energy, boundary faces included (2D:1026-1047, SL:1211-1227): the
    // force is read off one device step, F = (eta - eta') / (dt L_eta)
    {
        HostConfig hc = eta_config(dim);
        nicr_config& c = hc.c;
        c.dx = c.dy = c.dz = 1.;
        if (dim == 2) {
            c.eta_W = 1.;
            c.eta_kappa = 2.;
            c.eta_oxide_penalty = .7;
        } else {
            c.nx = 9;
            c.ny = 7;
            c.nz = 5;
            c.grain_diameter_grid = 6.;
            c.eta_W = 2.;
            c.eta_kappa = .3;
        }
        c.eta_L = .1;
        c.dt = 1.e-3;
        nicr_config_derive(&c);
        HostModel m(hc);
        for (int i = 1; i <= m.nx; ++i)
            for (int j = 1; j <= m.ny; ++j)
                for (int k = 1; k <= m.nz; ++k) {
                    const std::size_t q = m.at(i, j, k);
                    const double kk = dim == 3 ? k : 0.;
                    m.p[q] = .25 + .08 * std::sin(.31 * i + .27 * j + .2 * kk);
                    for (std::size_t g = 0; g < m.eta.size(); ++g)
                        m.eta[g][q] = .1 + .04 * std::cos(.2 * i + .3 * j + .7 * g + .5 * kk);
                }
        nicr_config dc = c;
        dc.n_eta = int(m.eta.size());
        Ctx d(dc);
        upload(m, d.h);
        nicr_step_report r{};
        check(nicr_step(d.h, 1, &r), "step");
        require(r.status == 0, "force step accepted");
        std::vector<double> e1(m.ncell);
        check(nicr_download_field(d.h, NICR_FIELD_ETA0, e1.data()), "download eta");
        Ctx en(dc);
        upload(m, en.h);
        const double eps = 1.e-5, vol = dim == 2 ? c.dx * c.dy : c.dx * c.dy * c.dz;
        double max_error = 0.;
        for (int i : {1, (m.nx + 1) / 2, m.nx})
            for (int j : {1, (m.ny + 1) / 2, m.ny})
                for (int k : {1, m.nz}) {
                    const std::size_t q = m.at(i, j, dim == 3 ? k : 1);
                    const double force = (m.eta[0][q] - e1[q]) / (c.dt * c.eta_L);
                    const double e0 = m.eta[0][q];
                    m.eta[0][q] = e0 + eps;
                    const long double plus = device_energy(m, en.h);
                    m.eta[0][q] = e0 - eps;
                    const long double minus = device_energy(m, en.h);
                    m.eta[0][q] = e0;
                    const double approx = double((plus - minus) / (2. * eps * vol));
                    const double err = std::abs(force - approx) / (1. + std::abs(force));
                    max_error = std::max(max_error, err);
                    require(err < (dim == 2 ? 1.e-7 : 2.e-7), "grain energy derivative equals the device eta force");
                    if (dim == 2) break;
                }
        std::cout << "ETA CHECK: largest scaled energy-derivative error of the device eta force=" << max_error << "\n";
    }
    // (4) closed grain-only energy is nonincreasing under the device step (2D:1064-1080, SL:1228-1230)
    {
        HostConfig hc = eta_config(dim);
        nicr_config& c = hc.c;
        c.dx = c.dy = c.dz = 1.;
        c.eta_W = 1.;
        c.eta_kappa = 2.;
        c.eta_L = 1.;
        c.dt = .005;
        if (dim == 3) c.dt = .002;
        nicr_config_derive(&c);
        HostModel m(hc);
        for (std::size_t q = 0; q < m.ncell; ++q) {
            m.o[q] = 0.;
            m.cr[q] = Phys::bulkCr;
            m.p[q] = 0.;
        }
        nicr_config dc = c;
        dc.n_eta = int(m.eta.size());
        Ctx d(dc);
        upload(m, d.h);
        Ctx en(dc);
        upload(m, en.h);
        long double previous = device_energy(m, en.h);
        const long double first = previous;
        for (int n = 0; n < 60; ++n) {
            nicr_step_report r{};
            check(nicr_step(d.h, 1, &r), "step");
            require(r.status == 0, "grain-only step accepted");
            for (std::size_t g = 0; g < m.eta.size(); ++g)
                check(nicr_download_field(d.h, NICR_FIELD_ETA0 + int(g), m.eta[g].data()), "download eta");
            const long double now = device_energy(m, en.h);
            require(now <= previous + 1.e-12L * (1.L + std::abs(previous)), "closed grain-only energy is nonincreasing");
            previous = now;
        }
        require(previous < first, "grain energy decreases");
        std::cout << "ETA CHECK: closed grain-only energy " << double(first) << " -> " << double(previous)
                  << " over 60 device steps\n";
    }
    std::cout << "PASS: device eta force is the energy variation; grain energy descends.\n";
}

int run(int argc, char** argv) {
    const int dim = NICR_CLI_DIM;
    PhaseTimes pt;
    try {
        HostConfig hc = parse_args(dim, argc, argv, std::cout);
        if (hc.self_test) {
            std::cout << "GPU SELF-TEST: device step of this build (" << backend_name() << ").\n";
            self_test(dim);
            eta_self_test(dim);
            return 0;
        }
        if (fs::exists(hc.out) && !fs::is_empty(hc.out)) throw std::runtime_error("Output directory is not empty; choose a new --out.");
        fs::create_directories(hc.out);
        // the 3D grain distances of the initial state on the GPU (bitwise the
        // host loops, which NICR_HOST_INIT=host selects)
        const char* init = std::getenv("NICR_HOST_INIT");
        hc.device_init = !(init && std::strcmp(init, "host") == 0);
        HostModel m(hc);
        pt.mark("init");
        const nicr_config& c = hc.c;
        nicr_config dc = c;
        dc.n_eta = int(m.eta.size());
        Device dev;
        check(nicr_create(&dc, &dev.ctx), "nicr_create");
        pt.mark("create");
        upload(m, dev.ctx);
        pt.mark("upload");
        const std::string backend = backend_name();
        m.metadata(backend);
        pt.mark("metadata");
        std::ofstream log, grain_diag, grain_history;
        log.open(hc.out + "/diagnostics.csv");
        if (!log) throw std::runtime_error("Cannot open diagnostics.csv");
        log << "step,time_s,sum_O,sum_Cr,mean_phi,O_boundary,O_seed,Cr_seed,O_Ck,Cr_Ck,O_clipping,Cr_clipping,"
               "O_budget_residual,Cr_budget_residual,nuclei,clipped_cells,min_O,max_O,min_Cr,max_Cr,min_Ni,min_phi,max_phi"
            << (dim == 3 ? ",mean_h_phi,oxide_volume_m3" : "") << "\n";
        if (m.has_eta()) {
            grain_diag.open(hc.out + "/grain_diagnostics.csv");
            grain_history.open(hc.out + "/grain_history.csv");
            if (!grain_diag || !grain_history) throw std::runtime_error("Cannot open grain diagnostics.");
            grain_diag << "step,time_s,min_eta,max_eta,min_sum_eta2,max_sum_eta2,mean_sum_eta2,grains_with_metal_cells,"
                          "unassigned_metal_cells,"
                       << (dim == 2 ? "integrated_active_GB_m2,max_GB_mask,grain_energy_per_depth\n"
                                    : "integrated_active_GB_m3,max_GB_mask,grain_energy_model_units\n");
            grain_history << "step,time_s,grain_id,dominant_metal_cells,"
                          << (dim == 2 ? "weighted_metal_area_cells,weighted_metal_area_m2\n"
                                       : "weighted_metal_volume_cells,weighted_metal_volume_m3\n");
        }
        pt.mark("metadata");
        download(m, dev.ctx);
        pt.mark("download");
        OutputThread writer;
        auto output = [&](int done) {
            writer.start([&m, &log, &grain_diag, &grain_history, &hc, done](OutputThread& w) {
                w.timed("log", [&] {
                    m.log_row(log, done);
                    log.flush();
                    if (!log) throw std::runtime_error("Diagnostics write failed.");
                });
                if (m.has_eta())
                    w.timed("grain_log", [&] {
                        m.grain_log(grain_diag, grain_history, done);
                        grain_diag.flush();
                        grain_history.flush();
                        if (!grain_diag || !grain_history) throw std::runtime_error("Grain diagnostics write failed.");
                    });
                w.timed("write", [&] { m.write_visualization(hc.out + "/fields_" + std::to_string(done)); });
            });
        };
        output(0);

        std::cout << "EXECUTION BACKEND: " << backend << " (HIP device step; no MPI).\n";
        std::cout << "MPI processes: 1 (X slabs)\n";
        if (dim == 3)
            std::cout << "Output: " << hc.output_format << ", " << (hc.output_format == "vtk" ? "ASCII Float64" : hc.output_precision)
                      << "; zlib level=" << hc.compression_level << "; frozen-eta cache=" << (c.frozen_eta_cache ? "on" : "off") << ".\n";
        std::cout << "Version " << (dim == 2 ? "mpi-eta-v1" : "mpi-eta-3d-xslab-optimized-v1") << "; grains=" << (c.grains ? "on" : "off")
                  << "; oxygen_diffusion=" << (c.oxygen_diffusion ? "on" : "off") << "; nucleation=" << hc.nucleation << "; "
                  << m.seeds.size() << " grain(s).\n";
        if (dim == 2)
            std::cout << "Box " << c.nx << " x " << c.ny << " cells; configured GB factors O=" << c.gb_factor_o << ", Cr=" << c.gb_factor_cr << ".\n";
        else
            std::cout << "Box " << c.nx << " x " << c.ny << " x " << c.nz << " cells; configured GB factors O=" << c.gb_factor_o
                      << ", Cr=" << c.gb_factor_cr << ".\n";
        if (c.grains)
            std::cout << "Voronoi random seed=" << c.random_seed << "; target grain diameter=" << c.grain_diameter_grid
                      << " grid = " << c.grain_diameter_grid * c.dx * 1.e6 << " um.\n";
        else
            std::cout << "Single crystal: grain_id=1, GB_mask=0; GB enhancement inactive.\n";
        if (m.has_eta()) {
            std::cout << "Evolving eta: " << m.eta.size() << " fields; W_eta=" << c.eta_W << ", kappa_eta=" << c.eta_K
                      << ", L_eta=" << c.eta_L << ", oxide penalty=" << c.eta_A << ".\n"
                      << "Eta 10-90 width=" << 2. * std::log(9.) * eta_ell(c) / c.dx
                      << " grids; GB conductance scale=" << c.eta_scale << ".\n"
                      << "NOTE: grain coefficients are provisional; dynamic coupling adds a grain-energy contribution to phi.\n";
            if (2. * std::log(9.) * eta_ell(c) / c.dx < 4.)
                std::cout << "WARNING: eta interface has fewer than 4 grid spacings across its 10-90 width.\n";
            if (dim == 2 && m.eta.size() > 64 && hc.eta_output == "all")
                std::cout << "NOTE: many eta output fields; --eta-output summary reduces file size.\n";
        }
        if (hc.nucleation == "center") {
            if (dim == 2)
                std::cout << "One initial hard oxide at box centre (" << 0.5 * c.nx * c.dx * 1.e6 << ", " << 0.5 * c.ny * c.dy * 1.e6
                          << ") um; radius=" << c.radius_grid << " grids; " << m.center_seed_cells
                          << " cells. Additional Ksp insertion is OFF.\n";
            else
                std::cout << "One initial hard oxide at box centre (" << 0.5 * c.nx * c.dx * 1.e6 << ", " << 0.5 * c.ny * c.dy * 1.e6
                          << ", " << 0.5 * c.nz * c.dz * 1.e6 << ") um; radius=" << c.radius_grid << " grids; "
                          << m.center_seed_cells << " cells. Additional Ksp insertion is OFF.\n";
        }
        if (!c.oxygen_diffusion)
            std::cout << "O diffusion and continuing O boundary supply are OFF; hard-seed assignments and optional Ck are unchanged.\n";
        if (hc.ck == "legacy") std::cout << "WARNING: legacy Ck is timestep dependent and nonconservative; retained only for comparison.\n";
        if (hc.bc == "mu" && c.oxygen_diffusion)
            std::cout << "WARNING: experimental mu BC uses a finite positive surface mobility closure (see README).\n";
        if (c.grains && ((c.oxygen_diffusion && c.gb_factor_o > 1) || c.gb_factor_cr > 1) && c.steps > 0)
            std::cout << (dim == 2 ? "WARNING: old timestep is NOT certified stable with enhanced GB mobility.\n"
                                   : "WARNING: 3D adds a third transport/curvature direction; check timestep convergence before production.\n");

        double next_nucleation = c.nucleation_period;
        int last_completed = 0;
        try {
            int n = 1;
            while (n <= c.steps) {
                // Ksp scan before step n when its physical time is reached (2D:1222)
                const double time = n * c.dt;
                if (hc.nucleation == "ksp" && time + 1.e-12 * c.nucleation_period >= next_nucleation) {
                    int cells = 0;
                    pt.mark("host");
                    check(nicr_nucleate(dev.ctx, &cells), "nucleate");
                    pt.mark("nucleate");
                    next_nucleation += c.nucleation_period;
                }
                // steps n..e-1 contain no event: no scan due, no output, no progress line
                int e = n;
                while (true) {
                    const int s = e;   // step s runs; events after it end the batch
                    if (s % 1000 == 0 || s == c.steps || s % c.output_every == 0) { ++e; break; }
                    const double tn = (s + 1) * c.dt;
                    if (s + 1 > c.steps) { ++e; break; }
                    if (hc.nucleation == "ksp" && tn + 1.e-12 * c.nucleation_period >= next_nucleation) { ++e; break; }
                    ++e;
                }
                const int batch = e - n;
                nicr_step_report r{};
                pt.mark("host");
                check(nicr_step(dev.ctx, batch, &r), "step");
                pt.mark("step");
                if (r.status != NICR_STEP_OK) {
                    last_completed = n - 1 + r.steps_done;
                    throw std::runtime_error(step_message(r, dim));
                }
                n = e;
                last_completed = n - 1;
                const int done = n - 1;
                if (done % 1000 == 0 || done == c.steps)
                    std::cout << "step " << done << "/" << c.steps << " (" << 100.0 * done / c.steps << "%), time=" << done * c.dt
                              << " s" << std::endl;
                if (done % c.output_every == 0 || done == c.steps) {
                    pt.mark("host");
                    writer.wait(pt);
                    pt.mark("output_wait");
                    download(m, dev.ctx);
                    pt.mark("download");
                    output(done);
                }
            }
            pt.mark("host");
            writer.wait(pt);
            pt.mark("output_wait");
        } catch (const std::exception& e) {
            writer.wait_quietly();
            {
                std::ofstream err(hc.out + "/STOPPED.txt");
                err << "Last accepted time step: " << last_completed << "\n" << e.what() << "\n";
            }
            download(m, dev.ctx, false);   // the fields only: no grain rows follow
            m.write_visualization(hc.out + "/state_at_failure");
            throw;
        }
        std::cout << "Finished " << last_completed << " steps. Output: " << hc.out << "\n";
        return 0;
    } catch (const EarlyExit&) {
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << "\n";
        return 1;
    }
}

}  // namespace

int main(int argc, char** argv) { return run(argc, argv); }
