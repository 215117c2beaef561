// Internal accessors shared by the engine's translation units.
#pragma once
#include <string>

#include "kernels.h"

namespace ice {
const DevSpace& space_dev(const ice_space* s);
struct HostMesh;
const HostMesh& space_mesh(const ice_space* s);
int space_batch(const ice_space* s);
bool space_coupled(const ice_space* s);  // periodic faces coupled as interior pairs
const PhysParams& disc_phys(const ice_disc* d);
int disc_flux(const ice_disc* d);
const ice_space* disc_space(const ice_disc* d);
int disc_neq(const ice_disc* d);
void set_error(const std::string& m);
int disc_residual(ice_disc* d, const double* u, double* r, hipStream_t st);
// Rebuild the u-independent tables that depend on the element vertices (the
// face geometry table and, with a source term, the source load) after the
// space's device vertex table changed; stream-ordered, allocates nothing.
int disc_refresh_geometry(ice_disc* d, hipStream_t st);
int disc_residual_in(ice_disc* d, const double* u, double* r, hipStream_t st, double* region, size_t cap,
                     size_t* need);
int disc_jacobian(ice_disc* d, const double* u, double* r, double* J, hipStream_t st);
int disc_dims(const ice_disc* d, int* nelem, int* nslots, int* bs);
const int* disc_jcols(const ice_disc* d);
// nel[3]: elements per direction of the mesh (not the total)
void disc_grid(const ice_disc* d, int* nel, int* neq, int* nbasis);
}  // namespace ice
