/* nicr_host.h -- host side of the drop-in programs (libnicr_host.so).
 *
 * The reference builds its initial state, parses its command line and writes
 * its output files inside the same single-file programs that step the model
 * (mpi2d/NiCr2D_grains.cpp, mpi3d-slab/NiCr3D_grains.cpp). This library holds
 * that host side, byte-compatible with the reference, for the drop-in programs
 * nicr2d/nicr3d and for tests; the step itself runs on the device (nicr.h).
 */
#ifndef NICR_HOST_H
#define NICR_HOST_H

#include <stddef.h>
#include "nicr.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* nicr_host_last_error(void);

/* Model::Model's state construction (mpi2d:209-228,244-327; mpi3d-slab:356-494):
 * Voronoi seeds, GB mask, eta fields, the concentration-BC layer and the centre
 * seed. bc, nucleation: the reference's option spellings. */
int nicr_host_create(const nicr_config* cfg, const char* bc, const char* nucleation, void** state);
void nicr_host_destroy(void* state);
int nicr_host_n_eta(void* state);
int nicr_host_n_seeds(void* state);
int nicr_host_center_seed_cells(void* state);
/* Field ids as nicr.h; 32 = initial grain id map. */
int nicr_host_get(void* state, int field, double* out);
/* Budget and initial sums {initialO, initialCr, ...} of the constructed state. */
int nicr_host_budget(void* state, nicr_budget* b, double initial[4]);
/* The reference's argument parsing (mpi2d:744-789, mpi3d-slab:1021-1081): fills
 * *out and writes the string options into buf as "bc\nck\nbounds\nnucleation\n
 * out\neta_output\noutput_format\noutput_precision\ncompression_level\nprecision".
 * Returns 0, 1 for --help/--version (their text in buf), -1 on an error (the
 * reference's message in buf). */
int nicr_host_parse(int dim, int argc, char** argv, nicr_config* out, char* buf, size_t len);
/* Writers driven from a given state (Model::log, grain_log, vtk, vti;
 * mpi2d:598-632,697-710; mpi3d-slab:233-334,827-872,963-977). what: 0 metadata
 * (path: a directory), 1 diagnostics row, 2 grain rows (path "diag|history"),
 * 3 legacy VTK, 4 VTI; options "format=...;precision=...;level=...;eta_output=...". */
int nicr_host_set_field(void* state, int field, const double* src);
int nicr_host_set_budget(void* state, const nicr_budget* b);
int nicr_host_write(void* state, int what, const char* path, int step, const char* options);
/* Seed coordinates (3 doubles per seed). */
int nicr_host_seeds(void* state, double* xyz);
/* n draws of std::uniform_real_distribution<double>(lo, hi) from
 * std::mt19937_64(seed), in order (the seeding of mpi3d-slab:425-427): the
 * composition noise of the benchmark states, drawn in the reference's loop order
 * when out[] is filled in storage order. No counterpart in the reference. */
int nicr_host_uniform_noise(unsigned long long seed, size_t n, double lo, double hi, double* out);

#ifdef __cplusplus
}
#endif
#endif /* NICR_HOST_H */
