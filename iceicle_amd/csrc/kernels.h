// Device-side views shared by the kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/iceicle_amd.h"

namespace ice {

template <class R>
struct PhysT;
using PhysParams = PhysT<double>;

// Everything a residual launch needs about the FE space, as raw device
// pointers (all owned by the ice_space handle).
struct DevSpace {
    int ndim, p, n;
    int nelem;
    int nfi;                 // interior faces (+ coupled periodic pairs), Q rows [0, nfi)
    int nfb;                 // boundary faces, Q rows [nfi, nfi + nfb)
    int norient;             // orientation codes per face number in fpt
    double period[3];
    const double* X;         // [nelem][2^ndim][ndim] element vertex coordinates
    const int4* fint;        // {elemL, elemR, info, 0}
    const int4* fbdy;        // {elemL, info, bcflag, 0}
    const int* efaces;       // [nelem][2*ndim]: (Q row << 4) | (side << 3) | orientation, -1 none
    const double* tab;       // xs[n+2] | L0[(n+2)n] | L1 | L2 | wgl[n]
    const float* tabf;       // FP32 copy of tab
    const signed char* fpt;  // [2*ndim][norient][nfq][ndim] S-indices of trace points
    const signed char* fvo;  // [2*ndim][4] face vertex order (element-local vertex ids)
    const int* jcols;        // [nelem][2*ndim+1] Jacobian block columns (-1: boundary face)
    const double* fgeo;      // [nfi + nfb][nfq][ndim + 2] unit normal, w sqrt(g), DDG h per face point
    const signed char* fperm;  // [2*ndim][norient][nfq] face point q -> the element's canonical point
    const signed char* finv;   // inverse of fperm
    const signed char* ftens;  // [2*ndim][norient][nfq] face point q -> its index in the element's
                               // tensor order of the face's tangential axes
    // Face colouring for the Jacobian assembly: no two faces of one colour
    // share an element, so a launch per colour adds into diagonal blocks
    // without atomics and in a fixed order.  jorder lists the interior faces
    // colour by colour, then the boundary faces colour by colour; colour c of
    // the interior faces is jorder[coff_i[c] .. coff_i[c + 1]), of the boundary
    // faces jorder[nfi + coff_b[c] .. nfi + coff_b[c + 1]).
    static constexpr int MAX_FACE_COLOURS = 16;
    const int* jorder;
    int ncol_i, ncol_b;
    int coff_i[MAX_FACE_COLOURS + 1], coff_b[MAX_FACE_COLOURS + 1];
};

// face info word: fnl | fnr << 3 | orient << 6 | bctype << 9 | pdim << 13
inline int pack_face_info(int fnl, int fnr, int orient, int bctype, int pdim) {
    return fnl | (fnr << 3) | (orient << 6) | (bctype << 9) | (pdim << 13);
}

// Per-discretization device data.
struct DiscDev {
    const ice_field* fields;  // [dirichlet x4 | neumann x4 | source]
    const double* bval;       // [nfb][nfq][neq] Dirichlet / Neumann values at boundary points, or null
    const double* srcload;    // [nelem][nbasis][neq] source load vector, or null
    const double* emu;        // [nelem] per-element viscosity (batched Burgers instances), or null
    int tile;                 // element-kernel workgroup: 0 = 256 threads, 1 = 128, 2 = 64 (whole elements)
    int tile_mixed;           // the same for the mixed-precision residual
    double* Q;                // [nfi + nfb][nfq][neq] weighted normal face fluxes
    double* Pic;              // [nfi + nfb][nfq][neq][ndim] DDGIC lifting vectors, or null
    double* T;                // [nelem][2*ndim][nfq][nv][neq] element face traces
    double* wave_part;        // per-block [max wave speed, max viscosity] when tracking, else null
    unsigned long long* wave; // wave-speed bookkeeping (bits of non-negative doubles), or null:
                              // [0] running max |dF/du| (Burgers lambda_max, burgers.hpp:89)
                              // [1] value at the last quadrature point (NS lambda_max, navier_stokes.hpp:930)
                              // [2] running max viscosity (NS visc_max, :964-966)
};

// Device-side state of an explicit time integration.
struct StepState {
    double dt;
    double time;
    long long itime;
    int done;
};

// Time-step and termination settings (FixedTimestep / CFLTimestep,
// TimestepTermination / TfinalTermination).
struct StepControl {
    int use_cfl;
    double dt_fixed, cfl, reflen, mu;
    int pmax, flux;
    int use_tfinal;
    double tfinal;
    long long ntime;
};

// Projection right-hand side b = sum_q f(x_q) phi_i w detJ (projection.hpp:83-110)
int projection_rhs(int ndim, int p, int neq, const DevSpace& S, const ice_field* f, double* b, hipStream_t st);
// Per-element squared L2 error terms (l2_error.hpp:37-94)
int l2_error_terms(int ndim, int p, int neq, const DevSpace& S, const ice_field* f, const double* U, double* out,
                   hipStream_t st);

int launch_minv(int ndim, int p, int neq, const DevSpace& S, const double* r, double* k, hipStream_t st);
int launch_minv_f32(int ndim, int p, int neq, const DevSpace& S, const double* r, float* k, hipStream_t st);
int launch_rk_combine_f32(size_t n, const float* base, int nk, const float* const* k, const double* c,
                          const double* dtp, const double* dte, size_t per, const StepState* stp, float* out,
                          hipStream_t st);
// k = M^-1 r and out = base + dt (c[0] kk[0] + .. + c[nprev - 1] kk[nprev - 1] + c[nprev] k), the
// step of launch_rk_combine (dte: one step per element), in one launch; k is stored only if
// store_k
int launch_minv_combine(int ndim, int p, int neq, const DevSpace& S, const double* r, double* k, const double* base,
                        int nprev, const double* const* kk, const double* c, const double* dtp, const double* dte,
                        const StepState* stp, double* out, bool store_k, hipStream_t st);
int launch_minv_combine_f32(int ndim, int p, int neq, const DevSpace& S, const double* r, float* k, const float* base,
                            int nprev, const float* const* kk, const double* c, const double* dtp, const double* dte,
                            const StepState* stp, float* out, bool store_k, hipStream_t st);
int launch_axpy_dt(size_t n, double coef, const double* dtp, const double* x, double* y, hipStream_t st);
int launch_axpy_edt(size_t n, double coef, const double* dte, size_t per, const StepState* stp, const double* x,
                    double* y, hipStream_t st);
int launch_rk_combine(size_t n, const double* base, int nk, const double* const* k, const double* c, const double* dtp,
                      const double* dte, size_t per, const StepState* stp, double* out, hipStream_t st);
int launch_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t st);
int launch_widen(size_t n, const float* x, double* y, hipStream_t st);  // y = (double)x
int launch_step_begin(StepState* s, unsigned long long* wave, const StepControl& c, hipStream_t st);
int launch_step_end(StepState* s, hipStream_t st);

int residual_fp64(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                  const double* U, double* R, hipStream_t st, hipEvent_t* ev = nullptr);
// FP32 interpolation, fluxes and face data with FP64 geometry and output;
// ndim 2 only (-2 otherwise)
int residual_mixed(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                   const double* U, double* R, hipStream_t st, hipEvent_t* ev = nullptr);
// The same with the coefficients read from an FP32 state (2D and 3D)
int residual_mixed_f32state(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                            const float* U, double* R, hipStream_t st);

// One-time tables of u-independent terms: boundary values at the boundary
// face quadrature points and the source load vector sum_q s(x_q) phi_i w detJ.
int setup_disc_tables(int ndim, int p, int neq, const DevSpace& S, const ice_field* fields, double* bval,
                      double* srcload, hipStream_t st);
// Load vector sum_q vals[e][q][v] phi_i w detJ (clamp: max(0, detJ)) from
// values at the volume quadrature points; physical coordinates of the
// boundary face quadrature points [nfb][nfq][ndim].
int load_from_values(int ndim, int p, int neq, const DevSpace& S, const double* vals, int clamp, double* out,
                     hipStream_t st);
int bface_coords(int ndim, int p, const DevSpace& S, double* out, hipStream_t st);

// Face geometry table of the space (see DevSpace::fgeo) and the size in
// doubles of a discretization's face trace buffer.
int face_geometry_table(int ndim, int p, const DevSpace& S, double* out, hipStream_t st);
size_t trace_buffer_size(int ndim, int p, int flux, int nelem);

// Finite-difference element-block Jacobian into J[nelem][2*ndim+1][bs][bs]
// (every block written; the same entries on every run); ndim <= 2, no DDGIC.
int jacobian_fp64(int ndim, int p, int flux, const DevSpace& S, const PhysParams& ph, const DiscDev& dd,
                  const double* U, double* J, size_t nvals, hipStream_t st);

// Block-sparse products y = J x (or y = b - J x when b is given), blocks
// [nelem][nslots][bs][bs] in T, vectors in V.
template <class T, class V>
int bsr_apply(int nelem, int nslots, int bs, const int* cols, const T* J, const V* x, const V* b, V* y,
              hipStream_t st);

}  // namespace ice
