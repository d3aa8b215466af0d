// Host-side helpers shared by the C-ABI translation units (npx_capi.cpp, npx_session.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "npx_kernels.h"
#include "npx_plan.h"

namespace npx {
namespace host {

// Sets the calling thread's npx_last_error() message and returns `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Geometry plan of (H, W, M, R), built on first use (host only).
int host_plan(int H, int W, int M, double R, const PlanHost** h);
// ... and its tables on the current device (uploaded on first use; not allowed under stream capture).
int tables_for(int H, int W, int M, double R, const PlanHost** h, const DevTables** t, hipStream_t st);

int pow_mode_of(double S);

// The plan's device tables as one contiguous blob (for callers that place them in memory of their own).
struct TableBlob {
    std::vector<uint8_t> bytes;
    size_t off[8];
    DevTables tables_in(void* mem) const;  // device view of the blob copied to `mem`
};
void pack_tables(const PlanHost& h, TableBlob* out);

// Launch shape of the fused kernels: wave slots per block, groups of slots, regions along x and y.
struct Shape {
    int nwaves, ngroups, nregx, nregy;
};

// Job fields shared by every entry point (stack, tables of one (H, W, M), launch shape); radii and sensitivities
// are added with add_radius / add_sensitivity.
FusedArgs job_args(const void* d_in, int dtype, int N, int H, int W, int M, const PlanHost& h, const DevTables* t,
                   const Shape& s, int iw);
void add_radius(FusedArgs& a, const PlanHost& h);
void add_sensitivity(FusedArgs& a, double S);

bool capturing(hipStream_t st);

// Frees the device caches of the other translation units (called by npx_release): radiality plans, eSRRF 3D
// tables, Fourier-interpolation FFT plans, shift-magnify tables.
void release_srrf();
void release_esrrf3d();
void release_fht();
void release_interp();

}  // namespace host
}  // namespace npx
