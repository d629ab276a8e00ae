// nicr_host.hpp -- host side of the drop-in programs (C++17, built with g++).
//
// The reference programs keep CLI parsing, the initial state built by the Model
// constructor, and the output writers in the same file as the time step
// (mpi2d/NiCr2D_grains.cpp, mpi3d-slab/NiCr3D_grains.cpp). Here they are host
// code around the device library: the initial state is computed on the host
// (libstdc++ <random> and glibc libm are part of its definition), uploaded once,
// and the device state is downloaded only at output time.
#pragma once
#include "../../../include/nicr.h"

#include <cmath>
#include <cstdint>
#include <iosfwd>
#include <stdexcept>
#include <string>
#include <vector>

namespace nicrh {

constexpr double PI = 3.14159265358979323846;

// Physics constants (2D:65-71, SL:73-79).
struct Phys {
    static constexpr double Vm = 1., W = 1e4, ko = 1e-9, kc = 7.5e-9, kp = 1.7e-10, L = 0.2;
    static constexpr double bmo = 2.21e-16, bmc = 3.5436e-20, boo = 2.21e-18, boc = 3.5436e-21;
    static constexpr double ckO = 2e-10, ckCr = 2e-10, ksp = 1.28e-7;
    static constexpr double bulkO = 0., bulkCr = 0.1986, layerO = 0.0023, layerCr = 0.182;
    static constexpr double oxideO = 0.56300203, oxideCr = 0.35677915;
};

// Everything the CLI accepts (Config, 2D:39-64 / SL:43-72) plus this build's
// own options, with the string options kept as strings for the writers.
struct HostConfig {
    nicr_config c{};
    std::string bc = "concentration", ck = "legacy", bounds = "stop", nucleation = "ksp";
    std::string out;
    std::string eta_output = "all";
    std::string output_format = "vti", output_precision = "float32";  // 3D only
    int compression_level = 1;                                         // 3D only
    std::string precision = "fp64";                                    // this build: fp64 | mixed
    bool self_test = false;
    int dim = 3;
    bool device_init = false;   // 3D grain distances of the initial state on the GPU (nicr_grain_distances)
    bool has_eta() const { return c.grain_evolution && c.grains; }
};

struct Seed { double x, y, z; };

struct LBudget {
    long double nucO = 0, nucCr = 0, ckO = 0, ckCr = 0, clipO = 0, clipCr = 0, bcO = 0, bcCr = 0;
    long long nuclei = 0, clipped = 0;
};

struct EarlyExit {};

// Parse argv as the reference does (2D:718-790 for dim 2, SL:986-1082 for dim 3):
// same flags, defaults, validation and messages. Throws std::runtime_error or EarlyExit.
HostConfig parse_args(int dim, int argc, char** argv, std::ostream& out);

// h(phi) (2D:152)
inline double h(double p) { return p * p * p * (10. - 15. * p + 6. * p * p); }
inline double minimum_image(double d, double period) { return d - period * std::round(d / period); }
double eta_ell(const nicr_config& c);
double reference_gb_width(const nicr_config& c);

// Host mirror of the Model state needed by initialisation and output. Fields
// hold interior cells in the reference's order (2D: Y fastest, 3D: Z fastest).
class HostModel {
public:
    HostConfig hc;
    int dim, nx, ny, nz;
    std::size_t ncell;
    std::vector<double> o, cr, p, gb, grain, gb_initial;
    std::vector<std::vector<double>> eta;
    std::vector<double> muO, muCr, mo, mc, eta_s2, eta_pairs;  // refreshed output arrays
    std::vector<Seed> seeds;
    std::vector<int> areas;  // areas (2D) or volumes (3D)
    LBudget b;
    long double initialO = 0, initialCr = 0;
    int center_seed_cells = 0;
    // The snapshot's per-cell grain-energy terms and order-free grain
    // statistics computed on the device (nicr_grain_terms); grain_log uses
    // them while grain_terms_valid, and its own loops otherwise.
    std::vector<double> grain_terms;
    double grain_extrema[5] = {0., 0., 0., 0., 0.};
    std::vector<int64_t> grain_counts;
    bool grain_terms_valid = false;

    explicit HostModel(const HostConfig& cfg);  // Model::Model (2D:209-228, SL:356-381)
    bool has_eta() const { return hc.has_eta(); }
    std::size_t at(int i, int j, int k = 1) const {   // 1-based reference indices
        return dim == 2 ? std::size_t(i - 1) * ny + (j - 1)
                        : (std::size_t(i - 1) * ny + (j - 1)) * nz + (k - 1);
    }

    // ---- output (all host-side, from downloaded device state) ----
    void metadata(const std::string& backend) const;                 // 2D:633-696, SL:874-962
    void log_row(std::ostream& f, int n) const;                       // 2D:697-710, SL:963-977
    void grain_log(std::ostream& f, std::ostream& history, int n) const;  // 2D:382-406, SL:566-590
    void write_vtk(const std::string& path) const;                    // 2D:598-632, SL:828-856
    void write_vti(const std::string& path) const;                    // SL:857-869 (+ writer SL:238-334)
    void write_visualization(const std::string& stem) const;          // SL:870-872

    int current_grain(std::size_t q) const;                           // 2D:360-366
    double grain_sum(std::size_t q) const;                            // SL:534-545
    long double grain_energy() const;                                 // 2D:367-381, SL:546-565

private:
    void voronoi();            // 2D:233-291, SL:407-467
    void initialize_eta();     // 2D:293-327, SL:469-494
    void initialize_center_oxide();  // 2D:517-542, SL:725-751
    double competitor_distance(std::size_t g, int owner, double x, double y, double z, double oy, double oz,
                               double rbest) const;   // SL:395-406
    // the same distance for two values of rbest, sharing each image's squared
    // distance and seed separation
    void competitor_distances(std::size_t g, int owner, double x, double y, double z, double oy, double oz,
                              double rbest_a, double rbest_b, double& da, double& db) const;
    void allocate_eta();
    std::vector<double> eta_norm;   // 3D: per-cell sum of the eta weights, between voronoi and initialize_eta
};

long double sum_field(const std::vector<double>& a);   // Model::sum (2D:229-232)

}  // namespace nicrh
