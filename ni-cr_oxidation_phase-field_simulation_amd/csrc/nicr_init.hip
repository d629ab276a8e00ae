// nicr_init.hip -- the per-cell distances of the initial 3D grain geometry on
// the device (nicr_grain_distances, include/nicr.h). Included by nicr_api.hip.
//
// Model::voronoi and initialize_eta (mpi3d-slab/NiCr3D_grains.cpp:395-406,
// 439-467, 469-494) give every cell its nearest seed (periodic Y/Z minimum
// image), the GB-mask distance to the nearest bisector of another seed's nine
// Y/Z images, and per seed the same distance measured from the owner image's
// offsets (the eta weights). That is O(cells x seeds x 9) binary64 work; the
// glibc exp() of the mask and of the weights stays on the host, which defines
// them. Every operation here is the host code's, in its order, with no
// contraction (-ffp-contract=off), pow(a, 2) as a * a (what g++ emits for it),
// IEEE division and square root, and std::min/std::max written as their
// comparisons, so each distance is bitwise the host's.

namespace nicr {

__device__ __forceinline__ double mi_period(double d, double period) { return d - period * round(d / period); }
__device__ __forceinline__ double min_std(double a, double b) { return (b < a) ? b : a; }   // std::min(a, b)
__device__ __forceinline__ double max_std(double a, double b) { return (a < b) ? b : a; }   // std::max(a, b)

// One thread per cell of planes [i0, i0 + planes). seeds: ng (x, y, z).
// owner, db: indexed by the cell's offset in the chunk; de: [ng][chunk cells].
__global__ __launch_bounds__(256) void k_grain_distances(int i0, int planes, int ny, int nz, int ng,
                                                         const double* __restrict__ seeds, int want_eta,
                                                         int* __restrict__ owner_out, double* __restrict__ db_out,
                                                         double* __restrict__ de_out) {
    const long long nyz = (long long)ny * nz, n = (long long)planes * nyz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int pi = (int)(t / nyz);
    const long long r = t - (long long)pi * nyz;
    const int j = (int)(r / nz) + 1, k = (int)(r - (long long)(j - 1) * nz) + 1, i = i0 + pi + 1;
    const double x = i - .5, y = j - .5, z = k - .5;
    const double pny = ny, pnz = nz;
    // nearest seed, first on ties (SL:443-450)
    int owner = 0;
    double rbest = __builtin_inf(), ownerY = 0., ownerZ = 0.;
    for (int g = 0; g < ng; ++g) {
        const double yy = mi_period(seeds[3 * g + 1] - y, pny), zz = mi_period(seeds[3 * g + 2] - z, pnz);
        const double ax = seeds[3 * g] - x;
        const double rr = ax * ax + yy * yy + zz * zz;
        if (rr < rbest) {
            rbest = rr;
            owner = g;
            ownerY = y + yy;
            ownerZ = z + zz;
        }
    }
    const double sox = seeds[3 * owner];
    const double ex = sox - x, ey = ownerY - y, ez = ownerZ - z;
    const double rbest_eta = ex * ex + ey * ey + ez * ez;   // initialize_eta's rbest (SL:482-483)
    double db = __builtin_inf();
    for (int g = 0; g < ng; ++g) {
        if (g == owner) {
            if (want_eta) de_out[(long long)g * n + t] = 0.;
            continue;
        }
        const double sx = seeds[3 * g], sy = seeds[3 * g + 1], sz = seeds[3 * g + 2];
        double da = __builtin_inf(), de = __builtin_inf();
        for (int im = -1; im <= 1; ++im)
            for (int jm = -1; jm <= 1; ++jm) {
                const double gy = sy + (double)(im * ny), gz = sz + (double)(jm * nz);
                const double a0 = sx - x, a1 = gy - y, a2 = gz - z;
                const double rs = a0 * a0 + a1 * a1 + a2 * a2;
                const double b0 = sx - sox, b1 = gy - ownerY, b2 = gz - ownerZ;
                const double sep = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
                if (sep > 1.e-14) {
                    da = min_std(da, max_std(0., (rs - rbest) / (2. * sep)));
                    de = min_std(de, max_std(0., (rs - rbest_eta) / (2. * sep)));
                }
            }
        db = min_std(db, da);
        if (want_eta) de_out[(long long)g * n + t] = de;
    }
    owner_out[t] = owner;
    db_out[t] = db;
}

}  // namespace nicr

extern "C" int nicr_grain_distances(int nx, int ny, int nz, int ng, const double* seeds, int32_t* owner, double* db,
                                    double* const* de) {
    using namespace nicr;
    if (nx < 1 || ny < 1 || nz < 1 || ng < 1 || !seeds || !owner || !db)
        return fail(NICR_E_INVALID, "nicr_grain_distances: bad argument");
    const long long nyz = (long long)ny * nz;
    const int want_eta = de != nullptr;
    // X-plane chunks with at most 2 GiB of per-seed distances on the device
    const long long per_plane = nyz * (8 + 4 + (want_eta ? 8LL * ng : 0));
    const int chunk = (int)std::max(1LL, std::min((long long)nx, (2LL << 30) / std::max(1LL, per_plane)));
    const long long cells = (long long)chunk * nyz;
    double* dseeds = nullptr;
    int* downer = nullptr;
    double* ddb = nullptr;
    double* dde = nullptr;
    hipStream_t s = nullptr;
    auto release = [&]() {
        if (s) hipStreamSynchronize(s);
        if (dseeds) hipFree(dseeds);
        if (downer) hipFree(downer);
        if (ddb) hipFree(ddb);
        if (dde) hipFree(dde);
        if (s) hipStreamDestroy(s);
    };
    auto check = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return NICR_OK;
        release();
        return fail(NICR_E_DEVICE, std::string("nicr_grain_distances: ") + what + ": " + hipGetErrorString(e));
    };
    int rc;
    if ((rc = check(hipStreamCreate(&s), "stream"))) return rc;
    if ((rc = check(hipMalloc((void**)&dseeds, (size_t)ng * 3 * 8), "alloc"))) return rc;
    if ((rc = check(hipMalloc((void**)&downer, (size_t)cells * 4), "alloc"))) return rc;
    if ((rc = check(hipMalloc((void**)&ddb, (size_t)cells * 8), "alloc"))) return rc;
    if (want_eta && (rc = check(hipMalloc((void**)&dde, (size_t)cells * 8 * ng), "alloc"))) return rc;
    if ((rc = check(hipMemcpyAsync(dseeds, seeds, (size_t)ng * 3 * 8, hipMemcpyHostToDevice, s), "copy"))) return rc;
    for (int i0 = 0; i0 < nx; i0 += chunk) {
        const int planes = std::min(chunk, nx - i0);
        const long long n = (long long)planes * nyz, off = (long long)i0 * nyz;
        hipLaunchKernelGGL(k_grain_distances, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, i0, planes, ny, nz,
                           ng, dseeds, want_eta, downer, ddb, dde);
        if ((rc = check(hipGetLastError(), "launch"))) return rc;
        if ((rc = check(hipMemcpyAsync(owner + off, downer, (size_t)n * 4, hipMemcpyDeviceToHost, s), "copy")))
            return rc;
        if ((rc = check(hipMemcpyAsync(db + off, ddb, (size_t)n * 8, hipMemcpyDeviceToHost, s), "copy"))) return rc;
        for (int g = 0; want_eta && g < ng; ++g)
            if ((rc = check(hipMemcpyAsync(de[g] + off, dde + (long long)g * n, (size_t)n * 8, hipMemcpyDeviceToHost, s),
                            "copy")))
                return rc;
        if ((rc = check(hipStreamSynchronize(s), "run"))) return rc;
    }
    release();
    return NICR_OK;
}
