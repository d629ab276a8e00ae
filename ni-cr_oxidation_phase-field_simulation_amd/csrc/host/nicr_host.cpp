// nicr_host.cpp -- drop-in host side: argument parsing, initial state, writers.
// See nicr_host.hpp. Output formats (key order, precision, array order) follow
// the reference byte for byte; they are interfaces of the programs.
#include "nicr_host.hpp"
#include "../host_threads.hpp"
#include "../../../include/nicr_host.h"

#include <zlib.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <locale>
#include <random>
#include <sstream>
#include <mutex>

namespace fs = std::filesystem;

namespace nicrh {

double reference_gb_width(const nicr_config& c) {
    return c.gb_fwhm_grid * c.dx * std::sqrt(PI) / (2. * std::sqrt(std::log(2.)));
}
double eta_ell(const nicr_config& c) { return std::sqrt(c.eta_K / (2. * c.eta_W)); }

long double sum_field(const std::vector<double>& a) {
    long double v = 0;
    for (double x : a) v += x;
    return v;
}

// ---------------------------------------------------------------------------
// argument parsing
// ---------------------------------------------------------------------------
static bool on_off(const std::string& value, const std::string& option) {
    if (value == "on") return true;
    if (value == "off") return false;
    throw std::runtime_error(option + " expects on or off (received: " + value + ")");
}

static void help(int dim, std::ostream& out) {
    if (dim == 2) {
        out << "Initialization only by default. Options:\n--steps N --dt s --nx N --ny N --dx m (sets dx=dy) --output-every N\n"
            << "--grains on|off (default on; off: uniform single crystal, no GB enhancement)\n"
            << "--grain-evolution on|off (default off; on: moving eta grains, dynamic GB transport and oxide coupling)\n"
            << "--eta-W value --eta-kappa value (0: auto width) --eta-L value --eta-oxide-penalty value (-1: eta-W)\n"
            << "--eta-output all|summary --max-eta-fields N (default 512; memory guard)\n"
            << "--oxygen-diffusion on|off (default on; off: no O transport or ongoing O boundary supply)\n"
            << "--left-edge-sites N (default 0; relocate N existing sites to partition inlet; ignored with grains off)\n"
            << "--grain-diameter-grid N --gb-width-grid N (mask FWHM)\n"
            << "--gb-factor-o N --gb-factor-cr N [--gb-factor N sets both; compatibility alias] --seed N\n"
            << "--nucleation ksp|center|off (default ksp; center: ONE initial seed; off: transport/grain growth only)\n"
            << "--radius-grid R (radius in grid spacings) --nucleation-period s (Ksp mode only)\n"
            << "--bc concentration|mu|closed --mu-res value --surface-c-ref value\n--ck legacy|off --bounds stop|legacy --out directory --self-test --version\n"
            << "Change --seed to change the Voronoi structure; repeat it with the same geometry/build to reproduce it.\n"
            << "O diffusion off does NOT disable hard-seed concentration assignments or legacy Ck.\n"
            << "With concentration BC, the initial left O/Cr layer is retained but O is not replenished when off.\n";
    } else {
        out << "Initialization only by default. Options:\n--steps N --dt s --nx N --ny N --nz N --dx m (sets dx=dy=dz) --output-every N\n"
            << "--grains on|off (default on; off: uniform single crystal, no GB enhancement)\n"
            << "--grain-evolution on|off (default off; on: moving eta grains, dynamic GB transport and oxide coupling)\n"
            << "--eta-W value --eta-kappa value (0: auto width) --eta-L value --eta-oxide-penalty value (-1: eta-W)\n"
            << "--eta-output all|summary (compatibility option; grain_sum replaces individual eta arrays)\n"
            << "--max-eta-fields N (default 512; memory guard)\n"
            << "--output-format vti|vtk (default vti; vtk: original ASCII double output)\n"
            << "--output-precision float32|float64 (default float32; VTI only)\n"
            << "--compression-level 0..9 (default 1, fast zlib; VTI only)\n"
            << "--frozen-eta-cache on|off (default on; off: rebuild every refresh for comparison)\n"
            << "Arrays: phi, conc_O, conc_Cr, mu_O, mu_Cr, grain_sum, GB_mask.\n"
            << "Output files are visualization snapshots, not full eta restart checkpoints.\n"
            << "grain_sum: raw weighted eta sum; without eta, initial_grain_id*(1-h(phi)).\n"
            << "GB_mask: existing 3D GB transport mask; eta overlap or fixed Voronoi, no extra phi mask.\n"
            << "--oxygen-diffusion on|off (default on; off: no O transport or ongoing O boundary supply)\n"
            << "--left-face-sites N (alias --left-edge-sites; relocate N existing sites over the YZ inlet)\n"
            << "--grain-diameter-grid N --gb-width-grid N (mask FWHM)\n"
            << "--gb-factor-o N --gb-factor-cr N [--gb-factor N sets both; compatibility alias] --seed N\n"
            << "--nucleation ksp|center|off (default ksp; center: ONE initial seed; off: transport/grain growth only)\n"
            << "--radius-grid R (radius in grid spacings) --nucleation-period s (Ksp mode only)\n"
            << "--bc concentration|mu|closed --mu-res value --surface-c-ref value\n--ck legacy|off --bounds stop|legacy --out directory --self-test --version\n"
            << "Change --seed to change the Voronoi structure; repeat it with the same 3D geometry/build to reproduce it.\n"
            << "O diffusion off does NOT disable hard-seed concentration assignments or legacy Ck.\n"
            << "With concentration BC, the initial left O/Cr layer is retained but O is not replenished when off.\n";
    }
    out << "GPU options of this build: --precision fp64|mixed (default fp64; mixed: binary32 stencils,\n"
        << "binary64 state, thresholds, inlet flux and budgets).\n";
}

HostConfig parse_args(int dim, int argc, char** argv, std::ostream& out) {
    HostConfig h;
    h.dim = dim;
    nicr_config_default(&h.c, dim);
    nicr_config& c = h.c;
    h.out = dim == 2 ? "nicr_grains_output" : "nicr3d_grains_output";
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        if (a == "--self-test") { h.self_test = true; continue; }
        if (a == "--version") { out << (dim == 2 ? "mpi-eta-v1" : "mpi-eta-3d-xslab-optimized-v1") << "\n"; throw EarlyExit{}; }
        if (a == "--help") { help(dim, out); throw EarlyExit{}; }
        if (k + 1 >= argc) throw std::runtime_error("Missing value for " + a);
        const std::string v = argv[++k];
        if (a == "--steps") c.steps = std::stoi(v);
        else if (a == "--dt") c.dt = std::stod(v);
        else if (a == "--nx") c.nx = std::stoi(v);
        else if (a == "--ny") c.ny = std::stoi(v);
        else if (dim == 3 && a == "--nz") c.nz = std::stoi(v);
        else if (a == "--dx") { c.dx = c.dy = std::stod(v); if (dim == 3) c.dz = c.dx; }
        else if (a == "--output-every") c.output_every = std::stoi(v);
        else if (a == "--grains") c.grains = on_off(v, a);
        else if (a == "--grain-evolution") c.grain_evolution = on_off(v, a);
        else if (a == "--eta-W") c.eta_W = std::stod(v);
        else if (a == "--eta-kappa") c.eta_kappa = std::stod(v);
        else if (a == "--eta-L") c.eta_L = std::stod(v);
        else if (a == "--eta-oxide-penalty") c.eta_oxide_penalty = std::stod(v);
        else if (a == "--eta-output") h.eta_output = v;
        else if (a == "--max-eta-fields") c.max_eta_fields = std::stoi(v);
        else if (dim == 3 && a == "--output-format") h.output_format = v;
        else if (dim == 3 && a == "--output-precision") h.output_precision = v;
        else if (dim == 3 && a == "--compression-level") {
            std::size_t used = 0;
            h.compression_level = std::stoi(v, &used);
            if (used != v.size()) throw std::runtime_error("--compression-level expects an integer 0..9.");
        } else if (dim == 3 && a == "--frozen-eta-cache") c.frozen_eta_cache = on_off(v, a);
        else if (a == "--oxygen-diffusion") c.oxygen_diffusion = on_off(v, a);
        else if (a == "--left-edge-sites" || (dim == 3 && a == "--left-face-sites")) {
            std::size_t used = 0;
            c.left_edge_sites = std::stoi(v, &used);
            if (used != v.size() || c.left_edge_sites < 0)
                throw std::runtime_error("--left-edge-sites expects a nonnegative integer.");
        } else if (a == "--grain-diameter-grid") c.grain_diameter_grid = std::stod(v);
        else if (a == "--gb-width-grid") c.gb_fwhm_grid = std::stod(v);
        else if (a == "--gb-factor-o") c.gb_factor_o = std::stod(v);
        else if (a == "--gb-factor-cr") c.gb_factor_cr = std::stod(v);
        else if (a == "--gb-factor") { c.gb_factor_o = std::stod(v); c.gb_factor_cr = c.gb_factor_o; }
        else if (a == "--nucleation") h.nucleation = v;
        else if (a == "--radius-grid") c.radius_grid = std::stod(v);
        else if (a == "--nucleation-period") c.nucleation_period = std::stod(v);
        else if (a == "--seed") c.random_seed = std::stoull(v);
        else if (a == "--bc") h.bc = v;
        else if (a == "--ck") h.ck = v;
        else if (a == "--bounds") h.bounds = v;
        else if (a == "--out") h.out = v;
        else if (a == "--mu-res") c.mu_res = std::stod(v);
        else if (a == "--surface-c-ref") c.surface_c_ref = std::stod(v);
        else if (a == "--precision") h.precision = v;
        else throw std::runtime_error("Unknown argument: " + a);
    }
    // validation (2D:770-789, SL:1055-1081)
    std::vector<double> finite = {c.dx, c.dy, c.dt, c.grain_diameter_grid, c.gb_fwhm_grid, c.gb_factor_o,
                                  c.gb_factor_cr, c.radius_grid, c.nucleation_period, c.mu_res, c.surface_c_ref};
    if (dim == 3) finite.insert(finite.begin() + 2, c.dz);
    for (double x : finite)
        if (!std::isfinite(x)) throw std::runtime_error("All numerical parameters must be finite.");
    const bool nz_bad = dim == 3 && (c.nz < 3 || c.nz > INT_MAX - 2 || c.dz <= 0);
    if (c.nx < 3 || c.ny < 3 || nz_bad || c.nx > INT_MAX - 2 || c.ny > INT_MAX - 2 || c.steps < 0 ||
        c.output_every < 1 || c.dt <= 0 || c.dx <= 0 || c.dy <= 0 || c.grain_diameter_grid <= 0 ||
        c.gb_fwhm_grid <= 0 || c.gb_factor_o < 1 || c.gb_factor_cr < 1 || c.radius_grid <= 0 ||
        (h.nucleation == "ksp" && c.nucleation_period < c.dt) || c.nucleation_period <= 0. || c.surface_c_ref <= 0)
        throw std::runtime_error("Invalid parameter bounds (Ksp nucleation period must be >= dt).");
    for (double v : {c.eta_W, c.eta_kappa, c.eta_L, c.eta_oxide_penalty})
        if (!std::isfinite(v)) throw std::runtime_error("Eta parameters must be finite.");
    if (c.eta_W <= 0. || c.eta_kappa < 0. || c.eta_L < 0. || (c.eta_oxide_penalty < 0. && c.eta_oxide_penalty != -1.) ||
        c.max_eta_fields < 1)
        throw std::runtime_error("Invalid eta parameter bounds.");
    if (dim == 3) {
        // "vtk-blocks" (no counterpart in mpi3d-slab): the legacy VTK array set of
        // the mpi3d-blocks program (BK:780-817)
        if (h.output_format != "vti" && h.output_format != "vtk" && h.output_format != "vtk-blocks")
            throw std::runtime_error("--output-format expects vti or vtk.");
        if (h.output_precision != "float32" && h.output_precision != "float64")
            throw std::runtime_error("--output-precision expects float32 or float64.");
        if (h.compression_level < 0 || h.compression_level > 9) throw std::runtime_error("--compression-level must lie in [0,9].");
        if (!std::isfinite(c.exclusion_gap_grid) || c.exclusion_gap_grid < 0. || !std::isfinite(c.radius_grid + c.exclusion_gap_grid))
            throw std::runtime_error("Nucleation exclusion length must be finite and nonnegative.");
    }
    if (h.eta_output != "all" && h.eta_output != "summary") throw std::runtime_error("--eta-output expects all or summary.");
    nicr_config_derive(&c);
    if (c.grain_evolution && c.grains && (!std::isfinite(c.eta_K) || !std::isfinite(c.eta_scale) || c.eta_K <= 0.))
        throw std::runtime_error("Derived eta coefficients are invalid.");
    if (h.nucleation != "ksp" && h.nucleation != "center" && h.nucleation != "off")
        throw std::runtime_error("--nucleation expects ksp, center, or off.");
    const int mn = dim == 2 ? std::min(c.nx, c.ny) : std::min({c.nx, c.ny, c.nz});
    if (h.nucleation == "center" && c.radius_grid >= 0.5 * (double(mn) - 1.))
        throw std::runtime_error(dim == 2 ? "Center seed is too large: --radius-grid must be less than (min(nx,ny)-1)/2."
                                          : "Center seed is too large: --radius-grid must be less than (min(nx,ny,nz)-1)/2.");
    if (dim == 3 && h.nucleation == "ksp" && c.radius_grid >= 0.5 * std::min(c.ny, c.nz))
        throw std::runtime_error("Ksp sphere radius must be less than min(ny,nz)/2 to avoid periodic self-overlap.");
    if (h.bc != "mu" && h.bc != "concentration" && h.bc != "closed") throw std::runtime_error("Invalid bc.");
    if (h.ck != "legacy" && h.ck != "off") throw std::runtime_error("Invalid Ck mode.");
    if (h.bounds != "stop" && h.bounds != "legacy") throw std::runtime_error("Invalid bounds mode.");
    if (h.precision != "fp64" && h.precision != "mixed") throw std::runtime_error("--precision expects fp64 or mixed.");
    c.bc = h.bc == "mu" ? NICR_BC_MU : (h.bc == "closed" ? NICR_BC_CLOSED : NICR_BC_CONCENTRATION);
    c.ck = h.ck == "off" ? NICR_CK_OFF : NICR_CK_LEGACY;
    c.bounds = h.bounds == "legacy" ? NICR_BOUNDS_LEGACY : NICR_BOUNDS_STOP;
    c.nucleation = h.nucleation == "center" ? NICR_NUC_CENTER : (h.nucleation == "off" ? NICR_NUC_OFF : NICR_NUC_KSP);
    c.precision = h.precision == "mixed" ? NICR_PRECISION_MIXED : NICR_PRECISION_FP64;
    if (dim == 2) c.nz = 1;
    return h;
}

// ---------------------------------------------------------------------------
// initial state (Model::Model)
// ---------------------------------------------------------------------------
HostModel::HostModel(const HostConfig& cfg) : hc(cfg) {
    const nicr_config& c = hc.c;
    dim = hc.dim;
    nx = c.nx;
    ny = c.ny;
    nz = dim == 3 ? c.nz : 1;
    if (dim == 3) {
        if (nx < 3 || ny < 3 || nz < 3) throw std::runtime_error("3D grid dimensions must lie in [3, INT_MAX-2].");
    }
    ncell = std::size_t(nx) * ny * nz;
    o.assign(ncell, 0.);
    cr.assign(ncell, 0.);
    p.assign(ncell, 0.);
    gb.assign(ncell, 0.);
    grain.assign(ncell, 0.);
    voronoi();
    // the Voronoi mask before eta takes over (2D's vtk and the vtk-blocks layout write it)
    if (dim == 2 || (c.grains && c.grain_evolution)) gb_initial = gb;
    const bool layer = hc.bc == "concentration";
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) {
                const std::size_t q = at(i, j, k);
                o[q] = Phys::bulkO;
                cr[q] = Phys::bulkCr;
                p[q] = 0.;
                if (layer && i == 1) {
                    o[q] = Phys::layerO;
                    cr[q] = Phys::layerCr;
                }
            }
    initialO = sum_field(o);
    initialCr = sum_field(cr);
    if (hc.nucleation == "center") initialize_center_oxide();
    if (has_eta()) initialize_eta();
}

using nicr_threads::parallel_chunks;

double HostModel::competitor_distance(std::size_t g, int owner, double x, double y, double z, double oy, double oz,
                                      double rbest) const {
    double db = std::numeric_limits<double>::infinity();
    for (int im = -1; im <= 1; ++im)
        for (int jm = -1; jm <= 1; ++jm) {
            const double gy = seeds[g].y + im * ny, gz = seeds[g].z + jm * nz;
            const double rs = std::pow(seeds[g].x - x, 2) + std::pow(gy - y, 2) + std::pow(gz - z, 2);
            const double sep = std::sqrt(std::pow(seeds[g].x - seeds[owner].x, 2) + std::pow(gy - oy, 2) +
                                         std::pow(gz - oz, 2));
            if (sep > 1.e-14) db = std::min(db, std::max(0., (rs - rbest) / (2. * sep)));
        }
    return db;
}

void HostModel::competitor_distances(std::size_t g, int owner, double x, double y, double z, double oy, double oz,
                                      double rbest_a, double rbest_b, double& da, double& db) const {
    da = db = std::numeric_limits<double>::infinity();
    for (int im = -1; im <= 1; ++im)
        for (int jm = -1; jm <= 1; ++jm) {
            const double gy = seeds[g].y + im * ny, gz = seeds[g].z + jm * nz;
            const double rs = std::pow(seeds[g].x - x, 2) + std::pow(gy - y, 2) + std::pow(gz - z, 2);
            const double sep = std::sqrt(std::pow(seeds[g].x - seeds[owner].x, 2) + std::pow(gy - oy, 2) +
                                         std::pow(gz - oz, 2));
            if (sep > 1.e-14) {
                da = std::min(da, std::max(0., (rs - rbest_a) / (2. * sep)));
                db = std::min(db, std::max(0., (rs - rbest_b) / (2. * sep)));
            }
        }
}

// eta fields zero-filled by host threads, whole fields per thread
void HostModel::allocate_eta() {
    eta.assign(seeds.size(), std::vector<double>());
    parallel_chunks(eta.size(), 1, [&](std::size_t g0, std::size_t g1) {
        for (std::size_t g = g0; g < g1; ++g) eta[g].assign(ncell, 0.);
    });
}

void HostModel::voronoi() {
    const nicr_config& c = hc.c;
    seeds.clear();
    areas.clear();
    if (!c.grains) {
        seeds.push_back({0.5 * nx, 0.5 * ny, 0.5 * nz});
        areas.push_back(nx * ny * nz);
        std::fill(grain.begin(), grain.end(), 1.);
        std::fill(gb.begin(), gb.end(), 0.);
        return;
    }
    int ng;
    if (dim == 2) {
        const double domain_area = double(nx) * ny;
        ng = std::max(1, int(std::llround(4. * domain_area / (PI * c.grain_diameter_grid * c.grain_diameter_grid))));
        if (ng > nx * ny) throw std::runtime_error("Too many Voronoi sites for this grid.");
    } else {
        const double volume = double(nx) * ny * nz;
        const double ng_real = 6. * volume / (PI * std::pow(c.grain_diameter_grid, 3));
        if (!std::isfinite(ng_real) || ng_real > volume + .5) throw std::runtime_error("Too many Voronoi sites for this grid.");
        ng = std::max(1, int(std::llround(ng_real)));
        if (ng > volume) throw std::runtime_error("Too many Voronoi sites for this grid.");
        if (has_eta() && ng > c.max_eta_fields)
            throw std::runtime_error("3D grain count exceeds --max-eta-fields; increase grain diameter or check memory before increasing the limit.");
    }
    std::mt19937_64 rng(c.random_seed);
    std::uniform_real_distribution<double> ux(0., double(nx)), uy(0., double(ny)), uz(0., double(nz));
    for (int g = 0; g < ng; ++g) {
        Seed s;
        s.x = ux(rng);
        s.y = uy(rng);
        s.z = dim == 3 ? uz(rng) : 0.;
        seeds.push_back(s);
    }
    const int les = c.left_edge_sites;
    if (dim == 2) {
        if (les < 0 || les > std::min(ng, ny))
            throw std::runtime_error("--left-edge-sites must be between 0 and min(area-based site count, ny).");
        if (les > 0) {
            std::uniform_real_distribution<double> unit(0., 1.);
            const int phase = std::min(ny - 1, int(unit(rng) * ny));
            for (int k = 0; k < les; ++k) {
                const int row = (phase + int((static_cast<long long>(k) * ny) / les)) % ny;
                const double sx = 0.05 + 0.4 * unit(rng);
                seeds[k] = {sx, row + 0.5, 0.};
            }
            for (int k = les; k < ng; ++k) seeds[k].x = 1.5 + (nx - 1.5) * (seeds[k].x / nx);
        }
    } else {
        if (les < 0 || les > std::min(ng, ny * nz))
            throw std::runtime_error("--left-face-sites must be between 0 and min(volume-based site count, ny*nz).");
        if (les > 0) {
            int rows = std::clamp(int(std::llround(std::sqrt(double(les) * ny / nz))), 1, std::min(les, ny));
            rows = std::max(rows, (les + nz - 1) / nz);
            std::uniform_real_distribution<double> unit(0., 1.);
            const int phaseY = std::min(ny - 1, int(unit(rng) * ny));
            const int phaseZ = std::min(nz - 1, int(unit(rng) * nz));
            int g = 0;
            for (int r = 0; r < rows; ++r) {
                const int cols = les / rows + (r < les % rows ? 1 : 0);
                const int yy = (phaseY + int(static_cast<long long>(r) * ny / rows)) % ny;
                for (int q = 0; q < cols; ++q) {
                    const int zz = (phaseZ + int(static_cast<long long>(q) * nz / cols)) % nz;
                    const double sx = 0.05 + 0.4 * unit(rng);
                    seeds[g++] = {sx, yy + 0.5, zz + 0.5};
                }
            }
            for (int k = les; k < ng; ++k) seeds[k].x = 1.5 + (nx - 1.5) * (seeds[k].x / nx);
        }
    }
    areas.assign(ng, 0);
    std::mutex areas_lock;
    // 3D with evolving grains: the eta weights (initialize_eta) come out of the
    // same competitor pass. Each weight is stored in its eta field and the
    // per-cell sum in eta_norm; initialize_eta then scales them by the metal
    // fraction once the center oxide is in place.
    const bool fuse_eta = dim == 3 && has_eta();
    const double ell_grid = eta_ell(c) / c.dx;
    if (fuse_eta) {
        allocate_eta();
        eta_norm.assign(ncell, 0.);
    }
    if (dim == 3 && hc.device_init) {
        // owners and distances from the device (bitwise the loop below); the
        // exp() of the mask and of the weights here, as below
        std::vector<double> sxyz(3 * std::size_t(ng));
        for (int g = 0; g < ng; ++g) {
            sxyz[3 * g] = seeds[g].x;
            sxyz[3 * g + 1] = seeds[g].y;
            sxyz[3 * g + 2] = seeds[g].z;
        }
        std::vector<int32_t> own(ncell);
        std::vector<double> dist(ncell);
        std::vector<double*> de;
        if (fuse_eta)
            for (auto& e : eta) de.push_back(e.data());
        if (nicr_grain_distances(nx, ny, nz, ng, sxyz.data(), own.data(), dist.data(), fuse_eta ? de.data() : nullptr) !=
            NICR_OK)
            throw std::runtime_error(std::string("Device grain geometry failed: ") + nicr_last_error());
        parallel_chunks(ncell, 1 << 14, [&](std::size_t q0, std::size_t q1) {
            std::vector<long long> my_areas(ng, 0);
            for (std::size_t q = q0; q < q1; ++q) {
                const int owner = own[q];
                grain[q] = owner + 1;
                ++my_areas[owner];
                const double db = dist[q];
                gb[q] = ng == 1 ? 0. : std::exp(-4. * std::log(2.) * db * db / (c.gb_fwhm_grid * c.gb_fwhm_grid));
                if (fuse_eta) {
                    double norm = 0.;
                    for (int g = 0; g < ng; ++g) {
                        const double w = g == owner ? 1. : std::exp(-eta[g][q] / ell_grid);
                        norm += w;
                        eta[g][q] = w;
                    }
                    eta_norm[q] = norm;
                }
            }
            std::lock_guard<std::mutex> guard(areas_lock);
            for (int g = 0; g < ng; ++g) areas[g] += my_areas[g];
        });
        return;
    }
    // X planes on host threads; each cell depends only on the seeds
    parallel_chunks(static_cast<std::size_t>(nx), 1, [&](std::size_t ia, std::size_t ib) {
    std::vector<long long> my_areas(ng, 0);
    for (int i = static_cast<int>(ia) + 1; i <= static_cast<int>(ib); ++i)
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) {
                const std::size_t q = at(i, j, k);
                const double x = i - .5, y = j - .5, z = k - .5;
                int owner = 0;
                double rbest = std::numeric_limits<double>::infinity(), ownerY = 0., ownerZ = 0.;
                if (dim == 2) {
                    for (int g = 0; g < ng; ++g) {
                        const double dy = minimum_image(seeds[g].y - y, ny);
                        const double r = (seeds[g].x - x) * (seeds[g].x - x) + dy * dy;
                        if (r < rbest) { rbest = r; owner = g; ownerY = y + dy; }
                    }
                } else {
                    for (int g = 0; g < ng; ++g) {
                        const double yy = minimum_image(seeds[g].y - y, ny), zz = minimum_image(seeds[g].z - z, nz);
                        const double rr = std::pow(seeds[g].x - x, 2) + yy * yy + zz * zz;
                        if (rr < rbest) { rbest = rr; owner = g; ownerY = y + yy; ownerZ = z + zz; }
                    }
                }
                grain[q] = owner + 1;
                ++my_areas[owner];
                double db = std::numeric_limits<double>::infinity();
                if (dim == 2) {
                    for (int g = 0; g < ng; ++g)
                        if (g != owner)
                            for (int im = -1; im <= 1; ++im) {
                                const double ym = seeds[g].y + im * ny;
                                const double rs = (seeds[g].x - x) * (seeds[g].x - x) + (ym - y) * (ym - y);
                                const double sep = std::hypot(seeds[g].x - seeds[owner].x, ym - ownerY);
                                if (sep > 1.e-14) db = std::min(db, std::max(0., (rs - rbest) / (2. * sep)));
                            }
                } else if (fuse_eta) {
                    // initialize_eta's rbest uses the owner image's offsets as oy - y, oz - z
                    const double rbest_eta =
                        std::pow(seeds[owner].x - x, 2) + std::pow(ownerY - y, 2) + std::pow(ownerZ - z, 2);
                    double norm = 0.;
                    for (int g = 0; g < ng; ++g) {
                        double w = 1.;
                        if (g != owner) {
                            double dm, de;
                            competitor_distances(g, owner, x, y, z, ownerY, ownerZ, rbest, rbest_eta, dm, de);
                            db = std::min(db, dm);
                            w = std::exp(-de / ell_grid);
                        }
                        norm += w;
                        eta[g][q] = w;
                    }
                    eta_norm[q] = norm;
                } else {
                    for (int g = 0; g < ng; ++g)
                        if (g != owner) db = std::min(db, competitor_distance(g, owner, x, y, z, ownerY, ownerZ, rbest));
                }
                gb[q] = ng == 1 ? 0. : std::exp(-4. * std::log(2.) * db * db / (c.gb_fwhm_grid * c.gb_fwhm_grid));
            }
    std::lock_guard<std::mutex> guard(areas_lock);
    for (int g = 0; g < ng; ++g) areas[g] += my_areas[g];
    });
}

void HostModel::initialize_eta() {
    const nicr_config& c = hc.c;
    if (seeds.size() > static_cast<std::size_t>(c.max_eta_fields))
        throw std::runtime_error(dim == 2 ? "Too many eta fields. Increase --max-eta-fields only after checking memory, or use --grain-evolution off."
                                          : "Too many eta fields; check --max-eta-fields and memory.");
    if (dim == 3) {
        // weights and their sums from the competitor pass in voronoi()
        parallel_chunks(ncell, 1 << 14, [&](std::size_t q0, std::size_t q1) {
            for (std::size_t q = q0; q < q1; ++q) {
                const double metal = 1. - h(p[q]), norm = eta_norm[q];
                for (std::size_t g = 0; g < eta.size(); ++g) eta[g][q] = metal * eta[g][q] / norm;
            }
        });
        std::vector<double>().swap(eta_norm);
        return;
    }
    allocate_eta();
    const double ell_grid = eta_ell(c) / c.dx;
    // X planes on host threads; each cell depends only on the seeds and its owner
    parallel_chunks(static_cast<std::size_t>(nx), 1, [&](std::size_t ia, std::size_t ib) {
    std::vector<double> weights(seeds.size());
    for (int i = static_cast<int>(ia) + 1; i <= static_cast<int>(ib); ++i)
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) {
                const std::size_t q = at(i, j, k);
                const int owner = int(grain[q]) - 1;
                const double x = i - .5, y = j - .5, z = k - .5;
                double norm = 0.;
                if (dim == 2) {
                    const double oy = y + minimum_image(seeds[owner].y - y, ny);
                    const double rbest = std::pow(seeds[owner].x - x, 2) + std::pow(oy - y, 2);
                    for (std::size_t g = 0; g < seeds.size(); ++g) {
                        if (int(g) == owner) { weights[g] = 1.; norm += 1.; continue; }
                        double distance = std::numeric_limits<double>::infinity();
                        for (int im = -1; im <= 1; ++im) {
                            const double gy = seeds[g].y + im * ny;
                            const double sep = std::hypot(seeds[g].x - seeds[owner].x, gy - oy);
                            const double r = std::pow(seeds[g].x - x, 2) + std::pow(gy - y, 2);
                            if (sep > 1.e-14) distance = std::min(distance, std::max(0., (r - rbest) / (2. * sep)));
                        }
                        weights[g] = std::exp(-distance / ell_grid);
                        norm += weights[g];
                    }
                } else {
                    const double oy = y + minimum_image(seeds[owner].y - y, ny);
                    const double oz = z + minimum_image(seeds[owner].z - z, nz);
                    const double rbest = std::pow(seeds[owner].x - x, 2) + std::pow(oy - y, 2) + std::pow(oz - z, 2);
                    for (std::size_t g = 0; g < seeds.size(); ++g) {
                        weights[g] = int(g) == owner ? 1.
                                                     : std::exp(-competitor_distance(g, owner, x, y, z, oy, oz, rbest) / ell_grid);
                        norm += weights[g];
                    }
                }
                const double metal = 1. - h(p[q]);
                for (std::size_t g = 0; g < eta.size(); ++g) eta[g][q] = metal * weights[g] / norm;
            }
    });
}

void HostModel::initialize_center_oxide() {
    const nicr_config& c = hc.c;
    const double x = 0.5 * (double(nx) + 1.), y = 0.5 * (double(ny) + 1.), z = 0.5 * (double(nz) + 1.);
    const double r = c.radius_grid;
    const int mn = dim == 2 ? std::min(nx, ny) : std::min({nx, ny, nz});
    if (r >= 0.5 * (double(mn) - 1.))
        throw std::runtime_error(dim == 2 ? "Center seed is too large: --radius-grid must be less than (min(nx,ny)-1)/2."
                                          : "Center seed is too large: --radius-grid must be less than (min(nx,ny,nz)-1)/2.");
    if (b.nuclei != 0) throw std::runtime_error("The center oxide has already been initialized.");
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k) {
                const double dx = i - x, dy = j - y, dz = k - z;
                const double d2 = dim == 2 ? dx * dx + dy * dy : dx * dx + dy * dy + dz * dz;
                if (d2 <= r * r) {
                    const std::size_t q = at(i, j, k);
                    b.nucO += Phys::oxideO - o[q];
                    b.nucCr += Phys::oxideCr - cr[q];
                    o[q] = Phys::oxideO;
                    cr[q] = Phys::oxideCr;
                    p[q] = 1.;
                    ++center_seed_cells;
                }
            }
    if (center_seed_cells == 0) throw std::runtime_error("Center seed covers no cell centres: increase --radius-grid.");
    ++b.nuclei;
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
int HostModel::current_grain(std::size_t q) const {
    if (p[q] >= .5) return 0;
    if (!has_eta()) return int(grain[q]);
    double best = 1.e-12;
    int id = 0;
    for (std::size_t g = 0; g < eta.size(); ++g)
        if (eta[g][q] > best) { best = eta[g][q]; id = int(g) + 1; }
    return id;
}

double HostModel::grain_sum(std::size_t q) const {
    if (has_eta()) {
        double value = 0.;
        for (std::size_t g = 0; g < eta.size(); ++g) value += double(g + 1) * eta[g][q];
        return value;
    }
    return grain[q] * (1. - h(p[q]));
}

static double local_energy(double S, double P, double phi, const nicr_config& c) {
    const double H = h(phi), d = S - (1. - H);
    return 0.25 * c.eta_W * d * d + c.eta_W * P + 0.5 * c.eta_A * H * S;
}

long double HostModel::grain_energy() const {
    if (!has_eta()) return 0.;
    if (grain_terms_valid && dim == 3) {
        long double local = 0.;
        for (std::size_t q = 0; q < ncell; ++q) local += grain_terms[q];
        return local;
    }
    const nicr_config& c = hc.c;
    // each cell's term (a double) on the host threads, then the long double
    // sum in the serial loop order, so the total is bitwise that of one thread
    std::vector<double> term(ncell);
    parallel_chunks(std::size_t(nx), 1, [&](std::size_t i0, std::size_t i1) {
        for (int i = int(i0) + 1; i <= int(i1); ++i)
            for (int j = 1; j <= ny; ++j)
                for (int k = 1; k <= nz; ++k) {
                    const std::size_t q = at(i, j, k);
                    double density;
                    if (dim == 2) {
                        density = local_energy(eta_s2[q], eta_pairs[q], p[q], c);
                    } else {
                        double S = 0., pairs = 0.;
                        for (const auto& e : eta) { const double v = e[q] * e[q]; pairs += S * v; S += v; }
                        density = local_energy(eta_s2[q], pairs, p[q], c);
                    }
                    const int jp = j == ny ? 1 : j + 1;
                    for (const auto& e : eta) {
                        if (i < nx) density += .5 * c.eta_K * std::pow((e[at(i + 1, j, k)] - e[q]) / c.dx, 2);
                        density += .5 * c.eta_K * std::pow((e[at(i, jp, k)] - e[q]) / c.dy, 2);
                        if (dim == 3) {
                            const int kp = k == nz ? 1 : k + 1;
                            density += .5 * c.eta_K * std::pow((e[at(i, j, kp)] - e[q]) / c.dz, 2);
                        }
                    }
                    term[q] = dim == 2 ? density * c.dx * c.dy : density * c.dx * c.dy * c.dz;
                }
    });
    long double local = 0.;
    for (std::size_t q = 0; q < ncell; ++q) local += term[q];
    return local;
}

void HostModel::grain_log(std::ostream& f, std::ostream& history, int n) const {
    if (!has_eta()) return;
    const nicr_config& c = hc.c;
    const std::size_t ng = eta.size();
    std::vector<long double> amount(ng, 0.);
    std::vector<int> cells(ng, 0);
    double lo = 1., hi = 0., slo = std::numeric_limits<double>::infinity(), shi = 0., gbmax = 0.;
    int unassigned = 0;
    std::vector<double> metal(ncell);
    // extrema and counts do not depend on the order: from the device, or
    // per-thread partials, merged
    std::mutex merge;
    const bool dev = grain_terms_valid && dim == 3;
    if (dev) {
        lo = grain_extrema[0];
        hi = grain_extrema[1];
        slo = grain_extrema[2];
        shi = grain_extrema[3];
        gbmax = grain_extrema[4];
        unassigned = int(grain_counts[0]);
        for (std::size_t g = 0; g < ng; ++g) cells[g] = int(grain_counts[g + 1]);
        parallel_chunks(ncell, 1 << 14, [&](std::size_t q0, std::size_t q1) {
            for (std::size_t q = q0; q < q1; ++q) metal[q] = 1. - h(p[q]);
        });
    } else
    parallel_chunks(ncell, 1 << 14, [&](std::size_t q0, std::size_t q1) {
        double tlo = 1., thi = 0., tslo = std::numeric_limits<double>::infinity(), tshi = 0., tgb = 0.;
        int tun = 0;
        std::vector<int> tcells(ng, 0);
        for (std::size_t q = q0; q < q1; ++q) {
            const double S = eta_s2[q];
            metal[q] = 1. - h(p[q]);
            tslo = std::min(tslo, S);
            tshi = std::max(tshi, S);
            tgb = std::max(tgb, gb[q]);
            const int id = current_grain(q);
            if (id) ++tcells[std::size_t(id - 1)];
            else if (p[q] < .5) ++tun;
            for (std::size_t g = 0; g < ng; ++g) {
                tlo = std::min(tlo, eta[g][q]);
                thi = std::max(thi, eta[g][q]);
            }
        }
        std::lock_guard<std::mutex> lock(merge);
        lo = std::min(lo, tlo);
        hi = std::max(hi, thi);
        slo = std::min(slo, tslo);
        shi = std::max(shi, tshi);
        gbmax = std::max(gbmax, tgb);
        unassigned += tun;
        for (std::size_t g = 0; g < ng; ++g) cells[g] += tcells[g];
    });
    // long double sums keep the serial cell order: one pass for the scalars,
    // and each thread owns whole grains for the per-grain amounts
    long double sumS = 0., mask = 0.;
    for (std::size_t q = 0; q < ncell; ++q) {
        sumS += eta_s2[q];
        mask += metal[q] * gb[q];
    }
    parallel_chunks(ng, 1, [&](std::size_t g0, std::size_t g1) {
        std::vector<long double> acc(g1 - g0, 0.);
        for (std::size_t q = 0; q < ncell; ++q) {
            const double S = eta_s2[q];
            if (S > 1.e-24)
                for (std::size_t g = g0; g < g1; ++g) {
                    const double v = eta[g][q];
                    acc[g - g0] += metal[q] * v * v / S;
                }
        }
        std::copy(acc.begin(), acc.end(), amount.begin() + std::ptrdiff_t(g0));
    });
    const auto alive = std::count_if(cells.begin(), cells.end(), [](int v) { return v > 0; });
    // long double products evaluated left to right, as the reference writes them
    auto per_volume = [&](long double v) { return dim == 2 ? v * c.dx * c.dy : v * c.dx * c.dy * c.dz; };
    f << std::setprecision(17) << n << "," << n * c.dt << "," << lo << "," << hi << "," << slo << "," << shi << ","
      << sumS / (dim == 2 ? (int)(nx * ny) : (int)(nx * ny * nz)) << "," << alive << "," << unassigned << ","
      << per_volume(mask) << "," << gbmax << "," << grain_energy() << "\n";
    for (std::size_t g = 0; g < eta.size(); ++g)
        history << std::setprecision(17) << n << "," << n * c.dt << "," << g + 1 << "," << cells[g] << "," << amount[g]
                << "," << per_volume(amount[g]) << "\n";
    if (!f || !history) throw std::runtime_error("Grain diagnostic write failed.");
}

void HostModel::log_row(std::ostream& f, int n) const {
    const nicr_config& c = hc.c;
    const long double so = sum_field(o), sc = sum_field(cr), po = sum_field(p);
    long double sumH = 0.;
    const long double ro = so - initialO - b.bcO - b.nucO - b.ckO - b.clipO;
    const long double rc = sc - initialCr - b.bcCr - b.nucCr - b.ckCr - b.clipCr;
    double omin = 1., omax = 0., crmin = 1., crmax = 0., nimin = 1., phimin = 1., phimax = 0.;
    for (std::size_t q = 0; q < ncell; ++q) {
        if (dim == 3) sumH += h(p[q]);
        omin = std::min(omin, o[q]);
        omax = std::max(omax, o[q]);
        crmin = std::min(crmin, cr[q]);
        crmax = std::max(crmax, cr[q]);
        nimin = std::min(nimin, 1. - o[q] - cr[q]);
        phimin = std::min(phimin, p[q]);
        phimax = std::max(phimax, p[q]);
    }
    const int cells = dim == 2 ? nx * ny : nx * ny * nz;
    f << std::setprecision(17) << n << "," << n * c.dt << "," << so << "," << sc << "," << po / cells << "," << b.bcO
      << "," << b.nucO << "," << b.nucCr << "," << b.ckO << "," << b.ckCr << "," << b.clipO << "," << b.clipCr << ","
      << ro << "," << rc << "," << b.nuclei << "," << b.clipped << "," << omin << "," << omax << "," << crmin << ","
      << crmax << "," << nimin << "," << phimin << "," << phimax;
    if (dim == 3) f << "," << sumH / cells << "," << sumH * c.dx * c.dy * c.dz;
    f << "\n";
}

// ---------------------------------------------------------------------------
// metadata: parameters.txt and grains.csv
// ---------------------------------------------------------------------------
void HostModel::metadata(const std::string& backend) const {
    const nicr_config& c = hc.c;
    std::ofstream f(hc.out + "/parameters.txt");
    const bool ev = has_eta();
    if (dim == 2) {
        f << std::setprecision(17)
          << (ev ? "STATUS: COUPLED ETA/OXIDATION MODEL; GRAIN PARAMETERS REQUIRE CALIBRATION\n"
                 : "STATUS: UNCALIBRATED FIXED-GRAIN TRANSPORT PROTOTYPE\n")
          << "nx=" << nx << "\nny=" << ny << "\ndx_m=" << c.dx << "\ndy_m=" << c.dy << "\ndt_s=" << c.dt
          << "\nsteps=" << c.steps << "\noxygen_bc=" << hc.bc << "\nCk=" << hc.ck << "\nbounds=" << hc.bounds
          << "\nmpi_processes=1\nmpi_decomposition=X slabs"
          << "\nversion=mpi-eta-v1\nnucleation_mode=" << hc.nucleation
          << "\ninitial_center_seed_cells=" << center_seed_cells << "\ngrain_structure=" << (c.grains ? "on" : "off")
          << "\noxygen_diffusion=" << (c.oxygen_diffusion ? "on" : "off")
          << "\noxygen_boundary_supply=" << (c.oxygen_diffusion ? hc.bc : std::string("disabled"))
          << "\nW=" << Phys::W << "\nkappa_phi=" << Phys::kp << "\nL_phi=" << Phys::L
          << "\nmu_res_model_units=" << c.mu_res << "\nsurface_mobility_concentration_reference=" << c.surface_c_ref
          << "\nleft_edge_sites_requested=" << c.left_edge_sites
          << "\nleft_edge_sites_active=" << (c.grains ? c.left_edge_sites : 0)
          << "\ngrain_diameter_target_grid=" << c.grain_diameter_grid
          << "\ngrain_diameter_target_m=" << c.grain_diameter_grid * c.dx << "\nvoronoi_sites=" << seeds.size()
          << "\ngb_mask_FWHM_grid=" << c.gb_fwhm_grid << "\ngb_mask_FWHM_m=" << c.gb_fwhm_grid * c.dx
          << "\ngb_mask_integral_effective_width_grid=" << c.gb_fwhm_grid * std::sqrt(PI) / (2. * std::sqrt(std::log(2.)))
          << "\ngb_mask_integral_effective_width_m=" << c.gb_fwhm_grid * c.dx * std::sqrt(PI) / (2. * std::sqrt(std::log(2.)))
          << "\ngb_peak_mobility_factor_O=" << c.gb_factor_o << "\ngb_peak_mobility_factor_Cr=" << c.gb_factor_cr
          << "\nrandom_seed=" << c.random_seed << "\nnucleus_geometric_radius_grid=" << c.radius_grid
          << "\nnucleus_geometric_radius_m=" << c.radius_grid * c.dx << "\nnucleation_period_s=" << c.nucleation_period
          << "\nNOTE: Vm=1 is retained, not evidence that energy/mobility units have been physically audited.\n"
          << "NOTE: no GB migration, GB energy, segregation, oxide GBs, or interphase enhancement.\n"
          << "NOTE: cell-centred grid; periodic Y; no-flux right; natural concentration/phi derivatives in X.\n";
        f << "execution_backend=" << backend << "\n";
        f << "grain_evolution_requested=" << (c.grain_evolution ? "on" : "off")
          << "\ngrain_evolution_active=" << (ev ? "on" : "off") << "\n";
        if (ev) {
            const double ell = eta_ell(c);
            f << "eta_fields=" << eta.size() << "\neta_W=" << c.eta_W << "\neta_kappa=" << c.eta_K << "\neta_L=" << c.eta_L
              << "\neta_oxide_penalty=" << c.eta_A << "\neta_width_10_90_grid=" << 2. * std::log(9.) * ell / c.dx
              << "\neta_mask_integrated_width_m=" << 8. * ell / 3. << "\neta_transport_scale=" << c.eta_scale
              << "\neta_planar_GB_energy_model_units=" << std::sqrt(2. * c.eta_K * c.eta_W) / 3.
              << "\neta_planar_geometric_mobility_model_units=" << 3. * c.eta_L * ell << "\neta_output=" << hc.eta_output
              << "\n"
              << "NOTE: New grain coefficients are provisional, not experimental Ni-Cr calibration.\n"
              << "NOTE: Dynamic eta modifies the GB pathways AND adds the reciprocal grain-energy force to phi.\n"
              << "NOTE: GB factors refer to the original integrated-width convention; GB_mask includes eta_transport_scale.\n";
        } else {
            f << "NOTE: fixed initial grain map and Gaussian GB mask.\n";
        }
        if (hc.nucleation == "center")
            f << "initial_oxide_center_x_m=" << 0.5 * nx * c.dx << "\ninitial_oxide_center_y_m=" << 0.5 * ny * c.dy
              << "\ninitial_oxide_center_i=" << 0.5 * (double(nx) + 1.) << "\ninitial_oxide_center_j="
              << 0.5 * (double(ny) + 1.)
              << "\nNOTE: one initial hard oxide only; no later Ksp insertion or replacement seed.\n";
        if (!c.grains) f << "NOTE: single grain; grain_id=1 and GB_mask=0; configured GB factors are inactive.\n";
        if (!c.oxygen_diffusion)
            f << "NOTE: O transport and continuing O boundary supply are off; initial inventory, hard seed assignments, optional Ck and bounds treatment remain.\n";
        std::ofstream g(hc.out + "/grains.csv");
        g << "grain_id,seed_x_grid,seed_y_grid,area_cells,equivalent_diameter_grid,equivalent_diameter_m\n";
        long double mean = 0;
        for (std::size_t k = 0; k < areas.size(); ++k) {
            double d = 2. * std::sqrt(areas[k] / PI);
            mean += d;
            g << std::setprecision(17) << k + 1 << "," << seeds[k].x << "," << seeds[k].y << "," << areas[k] << "," << d
              << "," << d * c.dx << "\n";
        }
        mean /= areas.size();
        f << "measured_arithmetic_mean_equivalent_diameter_grid=" << mean << "\n";
        if (!f || !g) throw std::runtime_error("Metadata write failed.");
        return;
    }
    // 3D (SL:874-962)
    const bool vtk = hc.output_format == "vtk";
    f << std::setprecision(17) << "model=NiCr_3D_oxidation_metal_grains\nversion=mpi-eta-3d-xslab-optimized-v1"
      << "\nnx=" << nx << "\nny=" << ny << "\nnz=" << nz << "\ndx_m=" << c.dx << "\ndy_m=" << c.dy << "\ndz_m=" << c.dz
      << "\ndt_s=" << c.dt << "\nsteps=" << c.steps << "\noutput_every=" << c.output_every
      << "\nvtk_fields=phi,conc_O,conc_Cr,mu_O,mu_Cr,grain_sum,GB_mask"
      << "\noutput_format=" << hc.output_format
      << "\noutput_precision=" << (vtk ? std::string("float64_ascii_17digits") : hc.output_precision)
      << "\noutput_compression=" << (hc.output_format == "vti" ? "zlib_raw_appended" : "none")
      << "\ncompression_level=" << hc.compression_level
      << "\nfrozen_eta_cache=" << (c.frozen_eta_cache ? "on" : "off")
      << "\nnucleation_search=bounded_periodic_neighborhoods_original_order_and_predicates"
      << "\nGB_mask_definition="
      << (!c.grains ? "zero;single_crystal"
                    : (ev ? "eta_transport_scale*min(1,16*sum_(a<b) eta_a^2*eta_b^2);no_extra_phi_mask"
                          : "exp(-4*ln(2)*distance_to_GB_grid^2/gb_fwhm_grid^2);fixed_Voronoi;zero_when_one_grain;no_extra_phi_mask"))
      << "\ngrain_sum_definition="
      << (ev ? "sum_(g=1..N) g*eta_g;unnormalized;no_extra_phi_mask"
             : "initial_grain_id*(1-h(phi));fixed_grain_fallback_no_eta_fields")
      << "\noxygen_bc=" << hc.bc << "\nCk=" << hc.ck << "\nbounds=" << hc.bounds
      << "\nmpi_processes=1\nmpi_decomposition=X_slabs_YZ_planes"
      << "\ntransverse_bc=periodic_Y_and_Z\nright_bc=no_flux"
      << "\nnucleation_mode=" << hc.nucleation << "\ninitial_center_seed_cells=" << center_seed_cells
      << "\ngrain_structure=" << (c.grains ? "on" : "off")
      << "\ngrain_evolution_requested=" << (c.grain_evolution ? "on" : "off")
      << "\ngrain_evolution_active=" << (ev ? "on" : "off")
      << "\noxygen_diffusion=" << (c.oxygen_diffusion ? "on" : "off")
      << "\noxygen_boundary_supply=" << (c.oxygen_diffusion ? hc.bc : std::string("disabled"))
      << "\nVm=" << Phys::Vm << "\nW=" << Phys::W << "\nkappa_phi=" << Phys::kp << "\nL_phi=" << Phys::L
      << "\nkappa_O=" << Phys::ko << "\nkappa_Cr=" << Phys::kc
      << "\nmatrix_mobility_prefactor_O=" << Phys::bmo << "\nmatrix_mobility_prefactor_Cr=" << Phys::bmc
      << "\noxide_mobility_prefactor_O=" << Phys::boo << "\noxide_mobility_prefactor_Cr=" << Phys::boc
      << "\nbulk_O=" << Phys::bulkO << "\nbulk_Cr=" << Phys::bulkCr
      << "\nlayer_O=" << Phys::layerO << "\nlayer_Cr=" << Phys::layerCr
      << "\noxide_O=" << Phys::oxideO << "\noxide_Cr=" << Phys::oxideCr
      << "\nCk_O=" << Phys::ckO << "\nCk_Cr=" << Phys::ckCr << "\nKsp=" << Phys::ksp
      << "\nmu_res_model_units=" << c.mu_res << "\nsurface_mobility_concentration_reference=" << c.surface_c_ref
      << "\nleft_face_sites_requested=" << c.left_edge_sites << "\nleft_face_sites_active=" << (c.grains ? c.left_edge_sites : 0)
      << "\ngrain_diameter_target_grid=" << c.grain_diameter_grid
      << "\ngrain_diameter_target_m=" << c.grain_diameter_grid * c.dx
      << "\ngrain_count_rule=round(6*nx*ny*nz/(pi*d_grid^3)),minimum_1"
      << "\nvoronoi_sites=" << seeds.size() << "\nrandom_seed=" << c.random_seed
      << "\ngb_mask_FWHM_grid=" << c.gb_fwhm_grid << "\ngb_mask_FWHM_m=" << c.gb_fwhm_grid * c.dx
      << "\ngb_mask_integral_effective_width_m=" << reference_gb_width(c)
      << "\ngb_peak_mobility_factor_O=" << c.gb_factor_o << "\ngb_peak_mobility_factor_Cr=" << c.gb_factor_cr
      << "\nnucleus_shape=sphere\nnucleus_geometric_radius_grid=" << c.radius_grid
      << "\nnucleus_geometric_radius_m=" << c.radius_grid * c.dx
      << "\nexclusion_gap_grid=" << c.exclusion_gap_grid << "\nnucleation_period_s=" << c.nucleation_period
      << "\neta_output=" << hc.eta_output << "\nmax_eta_fields=" << c.max_eta_fields << "\n";
    f << "execution_backend=" << backend << "\n";
    // The reference's storage estimate (halo arrays of its CPU Model): kept so
    // that the file stays comparable line by line.
    const double plane = double(nx + 2) * (ny + 2) * (nz + 2);
    const bool adv = ev && c.eta_L > 0.;
    double elements = 0.;
    elements += 10. * plane;                      // o, cr, p, on, crn, pn, muO, muCr, mo, mc
    elements += ev ? 0. : plane;                  // grain (released in eta mode)
    elements += plane;                            // gb
    elements += ev ? plane : 0.;                  // eta_s2
    elements += double(eta.size()) * plane * (adv ? 2. : 1.);
    const double bytes = sizeof(double) * elements;
    f << "rank0_field_storage_estimate_MiB=" << bytes / (1024. * 1024.)
      << "\nroot_additional_snapshot_field_storage_MiB=" << 0.
      << "\neta_next_fields_allocated=" << (adv ? eta.size() : 0)
      << "\ninitial_grain_map_retained=" << (!ev ? "yes" : "no")
      << "\neta_s2_storage_allocated=" << (ev ? "yes" : "no")
      << "\nfield_memory_estimates=steady_state_payload_including_halos;excluding_initialization_peak_and_MPI_overheads\n";
    if (ev) {
        const double ell = eta_ell(c);
        f << "eta_fields=" << eta.size() << "\neta_W=" << c.eta_W << "\neta_kappa=" << c.eta_K << "\neta_L=" << c.eta_L
          << "\neta_oxide_penalty=" << c.eta_A << "\neta_width_10_90_grid=" << 2. * std::log(9.) * ell / c.dx
          << "\neta_mask_integrated_width_m=" << 8. * ell / 3. << "\neta_transport_scale=" << c.eta_scale
          << "\neta_planar_GB_energy_model_units=" << std::sqrt(2. * c.eta_K * c.eta_W) / 3.
          << "\neta_planar_geometric_mobility_model_units=" << 3. * c.eta_L * ell << "\n";
    }
    if (hc.nucleation == "center")
        f << "initial_oxide_center_x_m=" << 0.5 * nx * c.dx << "\ninitial_oxide_center_y_m=" << 0.5 * ny * c.dy
          << "\ninitial_oxide_center_z_m=" << 0.5 * nz * c.dz << "\ninitial_oxide_center_i=" << 0.5 * (nx + 1.)
          << "\ninitial_oxide_center_j=" << 0.5 * (ny + 1.) << "\ninitial_oxide_center_k=" << 0.5 * (nz + 1.) << "\n";
    f << "mobility_note=prefactors_retained_from_uploaded_18wtpct_setup;local_M_includes_concentration_phase_and_GB_factors\n"
      << "calibration_note=eta_and_GB_transport_parameters_require_material_calibration\n";
    std::ofstream g(hc.out + "/grains.csv");
    g << "grain_id,seed_x_grid,seed_y_grid,seed_z_grid,volume_cells,volume_m3,equivalent_diameter_grid,equivalent_diameter_m\n";
    long double mean = 0.;
    for (std::size_t q = 0; q < areas.size(); ++q) {
        const double d = std::cbrt(6. * areas[q] / PI);
        mean += d;
        g << std::setprecision(17) << q + 1 << "," << seeds[q].x << "," << seeds[q].y << "," << seeds[q].z << ","
          << areas[q] << "," << areas[q] * c.dx * c.dy * c.dz << "," << d << "," << d * c.dx << "\n";
    }
    mean /= areas.size();
    f << "measured_arithmetic_mean_equivalent_diameter_grid=" << mean << "\n";
    if (!f || !g) throw std::runtime_error("Metadata write failed.");
}

// ---------------------------------------------------------------------------
// visualisation writers
// ---------------------------------------------------------------------------
// ASCII rows of one VTK array, formatted by worker threads with the stream's
// own locale, precision and flags, then written in order: rows(r, out) writes
// row r (of nrows) to out.
template <class RowFn>
static void vtk_rows(std::ostream& f, std::size_t nrows, RowFn rows) {
    const std::size_t batch = std::max<std::size_t>(1, 4096);
    for (std::size_t r0 = 0; r0 < nrows; r0 += batch * nicr_threads::host_threads()) {
        const std::size_t r1 = std::min(nrows, r0 + batch * nicr_threads::host_threads());
        const std::size_t nt = std::max<std::size_t>(1, (r1 - r0 + batch - 1) / batch);
        std::vector<std::string> piece(nt);
        parallel_chunks(nt, 1, [&](std::size_t a, std::size_t b) {
            for (std::size_t t = a; t < b; ++t) {
                std::ostringstream out;
                out.imbue(f.getloc());
                out.precision(f.precision());
                out.flags(f.flags());
                const std::size_t lo = r0 + t * batch, hi = std::min(r1, lo + batch);
                for (std::size_t r = lo; r < hi; ++r) rows(r, out);
                piece[t] = out.str();
            }
        });
        for (const auto& sp : piece) f.write(sp.data(), static_cast<std::streamsize>(sp.size()));
    }
}

void HostModel::write_vtk(const std::string& path) const {
    const nicr_config& c = hc.c;
    std::ofstream f(path);
    if (!f) throw std::runtime_error("Cannot write " + path);
    if (dim == 2) {
        f << std::setprecision(16) << "# vtk DataFile Version 3.0\n"
          << (has_eta() ? "NiCr evolving metal grains; coordinates in metres\n"
                        : "NiCr fixed-grain prototype; coordinates in metres\n")
          << "ASCII\nDATASET STRUCTURED_POINTS\n"
          << "DIMENSIONS " << nx << " " << ny << " 1\nORIGIN " << .5 * c.dx << " " << .5 * c.dy << " 0\nSPACING "
          << c.dx << " " << c.dy << " 1\nPOINT_DATA " << nx * ny << "\n";
        auto write = [&](const std::string& name, auto fn) {
            f << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
            vtk_rows(f, static_cast<std::size_t>(ny), [&](std::size_t r, std::ostream& out) {
                const int j = static_cast<int>(r) + 1;
                for (int i = 1; i <= nx; ++i) out << fn(at(i, j)) << " ";
                out << "\n";
            });
        };
        write("phi", [&](std::size_t q) { return p[q]; });
        write("conc_O", [&](std::size_t q) { return o[q]; });
        write("conc_Cr", [&](std::size_t q) { return cr[q]; });
        write("conc_Ni", [&](std::size_t q) { return 1. - o[q] - cr[q]; });
        write("grain_id", [&](std::size_t q) { return has_eta() ? double(current_grain(q)) : grain[q]; });
        write("GB_mask", [&](std::size_t q) { return gb[q]; });
        write("GB_active", [&](std::size_t q) { return (1. - h(p[q])) * gb[q]; });
        write("mu_O", [&](std::size_t q) { return muO[q]; });
        write("mu_Cr", [&](std::size_t q) { return muCr[q]; });
        write("M_O", [&](std::size_t q) { return mo[q]; });
        write("M_Cr", [&](std::size_t q) { return mc[q]; });
        if (has_eta()) {
            write("grain_id_initial", [&](std::size_t q) { return grain[q]; });
            write("GB_initial", [&](std::size_t q) { return gb_initial[q]; });
            write("GB_eta_indicator", [&](std::size_t q) { return std::min(1., 16. * eta_pairs[q]); });
            write("metal_fraction", [&](std::size_t q) { return 1. - h(p[q]); });
            write("eta_sum_sq", [&](std::size_t q) { return eta_s2[q]; });
            write("active_grain_id", [&](std::size_t q) { return double(current_grain(q)); });
            if (hc.eta_output == "all")
                for (std::size_t g = 0; g < eta.size(); ++g)
                    write("eta_" + std::to_string(g + 1), [&](std::size_t q) { return eta[g][q]; });
        }
    } else {
        f << std::setprecision(17) << "# vtk DataFile Version 3.0\nNiCr 3D oxidation and metal grains; metres\n"
          << "ASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS " << nx << " " << ny << " " << nz << "\nORIGIN "
          << .5 * c.dx << " " << .5 * c.dy << " " << .5 * c.dz << "\nSPACING " << c.dx << " " << c.dy << " " << c.dz
          << "\nPOINT_DATA " << nx * ny * nz << "\n";
        auto write = [&](const char* name, auto fn) {
            f << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
            vtk_rows(f, static_cast<std::size_t>(nz) * ny, [&](std::size_t r, std::ostream& out) {
                const int k = static_cast<int>(r / ny) + 1, j = static_cast<int>(r % ny) + 1;
                for (int i = 1; i <= nx; ++i) out << fn(at(i, j, k)) << " ";
                out << "\n";
            });
        };
        if (hc.output_format == "vtk-blocks") {
            // the array set and order of mpi3d-blocks (BK:780-817), that of the 2D program
            write("phi", [&](std::size_t q) { return p[q]; });
            write("conc_O", [&](std::size_t q) { return o[q]; });
            write("conc_Cr", [&](std::size_t q) { return cr[q]; });
            write("conc_Ni", [&](std::size_t q) { return 1. - o[q] - cr[q]; });
            write("grain_id", [&](std::size_t q) { return has_eta() ? double(current_grain(q)) : grain[q]; });
            write("GB_mask", [&](std::size_t q) { return gb[q]; });
            write("GB_active", [&](std::size_t q) { return (1. - h(p[q])) * gb[q]; });
            write("mu_O", [&](std::size_t q) { return muO[q]; });
            write("mu_Cr", [&](std::size_t q) { return muCr[q]; });
            write("M_O", [&](std::size_t q) { return mo[q]; });
            write("M_Cr", [&](std::size_t q) { return mc[q]; });
            if (has_eta()) {
                write("grain_id_initial", [&](std::size_t q) { return grain[q]; });
                write("GB_initial", [&](std::size_t q) { return gb_initial[q]; });
                write("GB_eta_indicator", [&](std::size_t q) { return std::min(1., 16. * eta_pairs[q]); });
                write("metal_fraction", [&](std::size_t q) { return 1. - h(p[q]); });
                write("eta_sum_sq", [&](std::size_t q) { return eta_s2[q]; });
                write("active_grain_id", [&](std::size_t q) { return double(current_grain(q)); });
                if (hc.eta_output == "all")
                    for (std::size_t g = 0; g < eta.size(); ++g) {
                        const std::string name = "eta_" + std::to_string(g + 1);
                        write(name.c_str(), [&](std::size_t q) { return eta[g][q]; });
                    }
            }
            if (!f) throw std::runtime_error("Output write failed: " + path);
            return;
        }
        write("phi", [&](std::size_t q) { return p[q]; });
        write("conc_O", [&](std::size_t q) { return o[q]; });
        write("conc_Cr", [&](std::size_t q) { return cr[q]; });
        write("mu_O", [&](std::size_t q) { return muO[q]; });
        write("mu_Cr", [&](std::size_t q) { return muCr[q]; });
        // grain_sum reads every eta field at the cell: evaluated in storage order
        std::vector<double> gsum(ncell);
        parallel_chunks(ncell, 1 << 14, [&](std::size_t q0, std::size_t q1) {
            for (std::size_t q = q0; q < q1; ++q) gsum[q] = grain_sum(q);
        });
        write("grain_sum", [&](std::size_t q) { return gsum[q]; });
        write("GB_mask", [&](std::size_t q) { return gb[q]; });
    }
    if (!f) throw std::runtime_error("Output write failed: " + path);
}

// VTK XML ImageData with raw appended zlib blocks (UInt64 headers), written to a
// temporary name and renamed into place (the layout of SL:238-334).
namespace {
class VtiWriter {
    static constexpr std::uint64_t kBlock = 65536;
    static constexpr std::array<const char*, 7> kNames = {"phi", "conc_O", "conc_Cr", "mu_O", "mu_Cr", "grain_sum", "GB_mask"};
    const HostModel& m;
    bool f64;
    int level;
    std::string final_path, temp_path;
    std::ofstream f;
    std::array<std::streampos, 7> offset_pos{};
    std::streampos appended_start;
    std::size_t next = 0;
    bool committed = false;

    void check() {
        if (!f) throw std::runtime_error("VTI write failed (check storage/quota): " + temp_path);
    }
    void put_u64(std::uint64_t v) {
        unsigned char b[8];
        for (int j = 0; j < 8; ++j) b[j] = static_cast<unsigned char>(v >> (8 * j));
        f.write(reinterpret_cast<const char*>(b), 8);
    }
    template <class T, class Fn>
    void blocks(Fn value) {
        const std::uint64_t n = std::uint64_t(m.nx) * m.ny * m.nz, total = n * sizeof(T);
        const std::uint64_t nb = (total + kBlock - 1) / kBlock;
        // the array in X-fastest order (the storage is Z-fastest), transposed by
        // worker threads in 64 x 64 (x, z) tiles of one y, so reads and writes
        // both run along 64 contiguous elements
        std::vector<T> all(static_cast<std::size_t>(n));
        constexpr int TT = 64;
        const int nkt = (m.nz + TT - 1) / TT;
        const std::size_t rowlen = static_cast<std::size_t>(m.nx), zstride = static_cast<std::size_t>(m.ny) * m.nx;
        parallel_chunks(static_cast<std::size_t>(m.ny) * nkt, 1, [&](std::size_t a, std::size_t b) {
            for (std::size_t t = a; t < b; ++t) {
                const int j = static_cast<int>(t / nkt) + 1, k0 = static_cast<int>(t % nkt) * TT + 1;
                const int k1 = std::min(m.nz, k0 + TT - 1);
                for (int i0 = 1; i0 <= m.nx; i0 += TT) {
                    const int i1 = std::min(m.nx, i0 + TT - 1);
                    for (int i = i0; i <= i1; ++i)
                        for (int k = k0; k <= k1; ++k) {
                            const double v = value(m.at(i, j, k));
                            if (std::isfinite(v) && std::abs(v) > std::numeric_limits<T>::max())
                                throw std::runtime_error(std::string("VTI ") + kNames[next] +
                                                         " exceeds output type range; use --output-precision float64.");
                            all[(k - 1) * zstride + static_cast<std::size_t>(j - 1) * rowlen + (i - 1)] =
                                static_cast<T>(v);
                        }
                }
            }
        });
        // independent zlib streams of kBlock bytes, compressed by worker threads
        std::vector<std::vector<unsigned char>> packed(static_cast<std::size_t>(nb));
        parallel_chunks(static_cast<std::size_t>(nb), 4, [&](std::size_t a, std::size_t b) {
            for (std::size_t q = a; q < b; ++q) {
                const std::uint64_t lo = q * kBlock, len = std::min<std::uint64_t>(kBlock, total - lo);
                packed[q].resize(static_cast<std::size_t>(compressBound(static_cast<uLong>(len))));
                uLongf dest = static_cast<uLongf>(packed[q].size());
                const int rc = compress2(packed[q].data(), &dest,
                                         reinterpret_cast<const Bytef*>(all.data()) + lo, static_cast<uLong>(len), level);
                if (rc != Z_OK) throw std::runtime_error("zlib compression failed: " + std::to_string(rc));
                packed[q].resize(static_cast<std::size_t>(dest));
            }
        });
        put_u64(nb);
        put_u64(kBlock);
        put_u64(total % kBlock);
        for (const auto& b : packed) put_u64(b.size());
        for (const auto& b : packed) f.write(reinterpret_cast<const char*>(b.data()), static_cast<std::streamsize>(b.size()));
        check();
    }

public:
    VtiWriter(const HostModel& model, const std::string& path)
        : m(model), f64(model.hc.output_precision == "float64"), level(model.hc.compression_level),
          final_path(path), temp_path(path + ".tmp") {
        const nicr_config& c = m.hc.c;
        try {
            f.open(temp_path, std::ios::binary | std::ios::trunc);
            if (!f) throw std::runtime_error("Cannot open " + temp_path);
            f.imbue(std::locale::classic());
            f << std::setprecision(17) << "<?xml version=\"1.0\"?>\n"
              << "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\" compressor=\"vtkZLibDataCompressor\">\n"
              << " <ImageData WholeExtent=\"0 " << m.nx - 1 << " 0 " << m.ny - 1 << " 0 " << m.nz - 1 << "\" Origin=\""
              << .5 * c.dx << " " << .5 * c.dy << " " << .5 * c.dz << "\" Spacing=\"" << c.dx << " " << c.dy << " "
              << c.dz << "\">\n"
              << "  <Piece Extent=\"0 " << m.nx - 1 << " 0 " << m.ny - 1 << " 0 " << m.nz - 1 << "\">\n"
              << "   <PointData Scalars=\"phi\">\n";
            for (std::size_t a = 0; a < kNames.size(); ++a) {
                f << "    <DataArray type=\"" << (f64 ? "Float64" : "Float32") << "\" Name=\"" << kNames[a]
                  << "\" NumberOfComponents=\"1\" format=\"appended\" offset=\"";
                offset_pos[a] = f.tellp();
                f << "00000000000000000000\"/>\n";
            }
            f << "   </PointData>\n   <CellData/>\n  </Piece>\n </ImageData>\n <AppendedData encoding=\"raw\">\n_";
            appended_start = f.tellp();
            check();
        } catch (...) {
            f.close();
            std::error_code ec;
            fs::remove(temp_path, ec);
            throw;
        }
    }
    ~VtiWriter() {
        if (!committed) {
            f.close();
            std::error_code ec;
            fs::remove(temp_path, ec);
        }
    }
    template <class Fn>
    void array(Fn value) {
        if (next >= kNames.size()) throw std::runtime_error("Too many VTI arrays.");
        const auto start = f.tellp();
        const auto offset = static_cast<std::uint64_t>(start - appended_start);
        f.seekp(offset_pos[next]);
        f << std::setw(20) << std::setfill('0') << offset << std::setfill(' ');
        f.seekp(start);
        check();
        if (f64) blocks<double>(value);
        else blocks<float>(value);
        ++next;
    }
    void finish() {
        if (next != kNames.size()) throw std::runtime_error("Missing VTI arrays.");
        f << "\n </AppendedData>\n</VTKFile>\n";
        f.flush();
        check();
        f.close();
        check();
        fs::rename(temp_path, final_path);
        committed = true;
    }
};
}  // namespace

void HostModel::write_vti(const std::string& path) const {
    VtiWriter w(*this, path);
    w.array([&](std::size_t q) { return p[q]; });
    w.array([&](std::size_t q) { return o[q]; });
    w.array([&](std::size_t q) { return cr[q]; });
    w.array([&](std::size_t q) { return muO[q]; });
    w.array([&](std::size_t q) { return muCr[q]; });
    // grain_sum reads every eta field at the cell: evaluated in storage order
    std::vector<double> gsum(ncell);
    parallel_chunks(ncell, 1 << 14, [&](std::size_t q0, std::size_t q1) {
        for (std::size_t q = q0; q < q1; ++q) gsum[q] = grain_sum(q);
    });
    w.array([&](std::size_t q) { return gsum[q]; });
    w.array([&](std::size_t q) { return gb[q]; });
    w.finish();
}

void HostModel::write_visualization(const std::string& stem) const {
    if (dim == 3 && hc.output_format == "vti") write_vti(stem + ".vti");
    else write_vtk(stem + ".vtk");
}

}  // namespace nicrh

// ---------------------------------------------------------------------------
// C entry points for tests: the host-built initial state (Model::Model)
// ---------------------------------------------------------------------------
extern "C" {
struct nicr_host_state { nicrh::HostModel* m; };
static thread_local std::string g_host_error;
const char* nicr_host_last_error(void) { return g_host_error.c_str(); }

int nicr_host_create(const nicr_config* cfg, const char* bc, const char* nucleation, void** out) {
    try {
        nicrh::HostConfig hc;
        hc.dim = cfg->dim;
        hc.c = *cfg;
        hc.bc = bc;
        hc.nucleation = nucleation;
        // the host loops by default; NICR_HOST_INIT=device takes the 3D grain
        // distances from the GPU
        const char* init = std::getenv("NICR_HOST_INIT");
        hc.device_init = init && std::strcmp(init, "device") == 0;
        nicr_config_derive(&hc.c);
        auto* s = new nicr_host_state{new nicrh::HostModel(hc)};
        *out = s;
        return 0;
    } catch (const std::exception& e) {
        g_host_error = e.what();
        return -1;
    }
}
void nicr_host_destroy(void* h) {
    auto* s = static_cast<nicr_host_state*>(h);
    if (s) { delete s->m; delete s; }
}
int nicr_host_n_eta(void* h) { return int(static_cast<nicr_host_state*>(h)->m->eta.size()); }
int nicr_host_n_seeds(void* h) { return int(static_cast<nicr_host_state*>(h)->m->seeds.size()); }
int nicr_host_center_seed_cells(void* h) { return static_cast<nicr_host_state*>(h)->m->center_seed_cells; }
// field ids as include/nicr.h; 32 = initial grain id map
int nicr_host_get(void* h, int field, double* out) {
    const nicrh::HostModel& m = *static_cast<nicr_host_state*>(h)->m;
    const std::vector<double>* f = nullptr;
    switch (field) {
    case NICR_FIELD_O: f = &m.o; break;
    case NICR_FIELD_CR: f = &m.cr; break;
    case NICR_FIELD_PHI: f = &m.p; break;
    case NICR_FIELD_GB: f = &m.gb; break;
    case 32: f = &m.grain; break;
    default:
        if (field >= NICR_FIELD_ETA0 && field < NICR_FIELD_ETA0 + int(m.eta.size())) f = &m.eta[field - NICR_FIELD_ETA0];
    }
    if (!f) { g_host_error = "unknown field"; return -1; }
    std::memcpy(out, f->data(), f->size() * sizeof(double));
    return 0;
}
int nicr_host_budget(void* h, nicr_budget* b, double initial[4]) {
    const nicrh::HostModel& m = *static_cast<nicr_host_state*>(h)->m;
    auto split = [](long double v, double o[2]) { o[0] = double(v); o[1] = double(v - (long double)o[0]); };
    std::memset(b, 0, sizeof(*b));
    split(m.b.nucO, b->nucO);
    split(m.b.nucCr, b->nucCr);
    b->nuclei = m.b.nuclei;
    split(m.initialO, initial);
    split(m.initialCr, initial + 2);
    return 0;
}
// The CLI parser (parse_args) for tests: fills *out and writes the string options
// as "bc\nck\nbounds\nnucleation\nout\neta_output\noutput_format\noutput_precision\n
// compression_level\nprecision" into buf. Returns 0, 1 for an early exit
// (--help/--version; their text goes to buf), -1 on a parse error.
int nicr_host_parse(int dim, int argc, char** argv, nicr_config* out, char* buf, size_t len) {
    std::ostringstream text;
    try {
        const nicrh::HostConfig h = nicrh::parse_args(dim, argc, argv, text);
        *out = h.c;
        std::ostringstream s;
        s << h.bc << "\n" << h.ck << "\n" << h.bounds << "\n" << h.nucleation << "\n" << h.out << "\n" << h.eta_output
          << "\n" << h.output_format << "\n" << h.output_precision << "\n" << h.compression_level << "\n" << h.precision
          << "\n" << (h.self_test ? 1 : 0);
        std::snprintf(buf, len, "%s", s.str().c_str());
        return 0;
    } catch (const nicrh::EarlyExit&) {
        std::snprintf(buf, len, "%s", text.str().c_str());
        return 1;
    } catch (const std::exception& e) {
        g_host_error = e.what();
        return -1;
    }
}
// Writers driven from externally supplied state (tests feed oracle fields).
int nicr_host_set_field(void* h, int field, const double* src) {
    nicrh::HostModel& m = *static_cast<nicr_host_state*>(h)->m;
    std::vector<double>* f = nullptr;
    switch (field) {
    case NICR_FIELD_O: f = &m.o; break;
    case NICR_FIELD_CR: f = &m.cr; break;
    case NICR_FIELD_PHI: f = &m.p; break;
    case NICR_FIELD_GB: f = &m.gb; break;
    case NICR_FIELD_MU_O: f = &m.muO; break;
    case NICR_FIELD_MU_CR: f = &m.muCr; break;
    case NICR_FIELD_M_O: f = &m.mo; break;
    case NICR_FIELD_M_CR: f = &m.mc; break;
    case NICR_FIELD_ETA_S2: f = &m.eta_s2; break;
    case NICR_FIELD_ETA_PAIRS: f = &m.eta_pairs; break;
    default:
        if (field >= NICR_FIELD_ETA0 && field < NICR_FIELD_ETA0 + int(m.eta.size())) f = &m.eta[field - NICR_FIELD_ETA0];
    }
    if (!f) { g_host_error = "unknown field"; return -1; }
    f->assign(src, src + m.ncell);
    return 0;
}
int nicr_host_set_budget(void* h, const nicr_budget* b) {
    nicrh::HostModel& m = *static_cast<nicr_host_state*>(h)->m;
    auto j = [](const double v[2]) { return (long double)v[0] + (long double)v[1]; };
    m.b.nucO = j(b->nucO); m.b.nucCr = j(b->nucCr); m.b.ckO = j(b->ckO); m.b.ckCr = j(b->ckCr);
    m.b.clipO = j(b->clipO); m.b.clipCr = j(b->clipCr); m.b.bcO = j(b->bcO); m.b.bcCr = j(b->bcCr);
    m.b.nuclei = b->nuclei; m.b.clipped = b->clipped;
    return 0;
}
// what: 0 metadata into `path` (a directory), 1 diagnostics row, 2 grain rows
// (path = "diag|history"), 3 legacy VTK, 4 VTI. Rows are appended.
int nicr_host_write(void* h, int what, const char* path, int step, const char* options) {
    nicrh::HostModel& m = *static_cast<nicr_host_state*>(h)->m;
    try {
        if (options && *options) {
            // "format=vti;precision=float32;level=1;eta_output=all"
            std::string s(options), item;
            std::istringstream in(s);
            while (std::getline(in, item, ';')) {
                const auto eq = item.find('=');
                if (eq == std::string::npos) continue;
                const std::string k = item.substr(0, eq), v = item.substr(eq + 1);
                if (k == "format") m.hc.output_format = v;
                else if (k == "precision") m.hc.output_precision = v;
                else if (k == "level") m.hc.compression_level = std::stoi(v);
                else if (k == "eta_output") m.hc.eta_output = v;
                else if (k == "ck") m.hc.ck = v;
                else if (k == "bounds") m.hc.bounds = v;
            }
        }
        // the writers read the refreshed arrays (mu, M, eta sums) that the caller
        // uploads after a refresh; a missing one is an error, not a stray read
        auto need = [&](const std::vector<double>& v, const char* name) {
            if (v.size() != m.ncell)
                throw std::runtime_error(std::string("nicr_host_write: ") + name +
                                         " is not set (upload the refreshed field first)");
        };
        if (what == 3 || what == 4) {
            need(m.muO, "mu_O");
            need(m.muCr, "mu_Cr");
        }
        const bool full_vtk = what == 3 && (m.dim == 2 || m.hc.output_format == "vtk-blocks");
        if (full_vtk) {
            need(m.mo, "M_O");
            need(m.mc, "M_Cr");
        }
        if ((what == 2 || full_vtk) && m.has_eta()) {
            need(m.eta_s2, "eta_s2");
            if (full_vtk) need(m.eta_pairs, "eta_pairs");
        }
        if (what == 0) {
            m.hc.out = path;
            m.metadata("single_process_verification_shim");
        } else if (what == 1) {
            std::ofstream f(path, std::ios::app);
            m.log_row(f, step);
        } else if (what == 2) {
            const std::string s(path);
            const auto bar = s.find('|');
            std::ofstream f(s.substr(0, bar), std::ios::app), g(s.substr(bar + 1), std::ios::app);
            m.grain_log(f, g, step);
        } else if (what == 3) {
            m.write_vtk(path);
        } else if (what == 4) {
            m.write_vti(path);
        } else {
            g_host_error = "unknown writer";
            return -1;
        }
        return 0;
    } catch (const std::exception& e) {
        g_host_error = e.what();
        return -1;
    }
}
int nicr_host_seeds(void* h, double* xyz) {
    const nicrh::HostModel& m = *static_cast<nicr_host_state*>(h)->m;
    std::size_t q = 0;
    for (const auto& s : m.seeds) { xyz[q++] = s.x; xyz[q++] = s.y; xyz[q++] = s.z; }
    return 0;
}
// Seeded composition noise of the benchmark workloads: n draws of
// std::uniform_real_distribution<double>(lo, hi) from std::mt19937_64(seed), in
// order. The reference seeds its Voronoi points this way (mpi3d-slab:425-427);
// filling out[] in storage order draws in the reference's loop order (i, then j,
// then k), so the same noise can be rebuilt by any program linked against the
// same C++ library.
int nicr_host_uniform_noise(unsigned long long seed, size_t n, double lo, double hi, double* out) {
    if (!out && n) {
        g_host_error = "nicr_host_uniform_noise: null output";
        return -1;
    }
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> u(lo, hi);
    for (size_t q = 0; q < n; ++q) out[q] = u(gen);
    return 0;
}
}
