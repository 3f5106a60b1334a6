// Internal definitions of the Poisson band (msbg_poisson) shared by the
// relaxation / solve unit and the learned-correction unit.
#pragma once

#include "internal.hpp"

namespace msbg {

// One listed value block: its slot, voxel origin and the slots of its six
// face neighbours (-x, +x, -y, +y, -z, +z; unused where the face is on the
// domain boundary).
struct PBlock {
    int32_t slot, ox, oy, oz;
    int32_t nb[6];
};

struct Dom {
    int32_t sx, sy, sz;
};

// Per-block kernels of 256 threads; the cell-wise ones (correction, dots) take
// kSlices work-groups per block, 4 z-slices each, 16 x 16 threads.
constexpr int kThreads = 256;
constexpr int kSlices = 4;

}  // namespace msbg

struct msbg_poisson {
    msbg::Dom dom{};
    int64_t n_slots = 0, n_active = 0, n_cells = 0;
    int64_t nx = 0, ny = 0, nz = 0, n_blocks = 0;
    int n_wg = 0;                     // work-groups of the kSlices-per-block kernels
    int zcol[3] = {1, 1, 1};          // column depth of the relaxation kernels, per storage
    msbg::PBlock *d_blocks = nullptr;
    int32_t *d_b2s = nullptr;        // copy of the layout grid's bid -> slot table
    uint8_t *d_listed = nullptr;     // per block: 1 if listed
    double2 *d_partial = nullptr;
    double *d_norm = nullptr;   // [2]
    float *d_scale = nullptr;   // [2]: 2^k ~ rms(r), 2^-k
    double *h_norm = nullptr;   // pinned [2]
    float *d_r = nullptr;       // FP32 residual, n_slots * 4096
    void *d_e = nullptr, *d_rt = nullptr;  // inner fields (2 or 4 B elements)
    int inner_bytes = 0;
    bool e_zeroed = false;
    // learned correction scratch
    float *d_delta = nullptr, *d_xtry = nullptr;
    double *d_dots = nullptr;  // [2]: r . A delta, |A delta|^2
    // graph of one check period, keyed by its arguments
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    int g_mode = -1, g_storage = -1, g_sweeps = -1;
    float g_omega = 0.0f;
    const void *g_x = nullptr, *g_b = nullptr;
    // graph of a run of Jacobi sweeps, keyed by its arguments
    hipGraphExec_t jexec = nullptr;
    int j_storage = -1, j_sweeps = -1;
    float j_omega = 0.0f;
    const void *j_x = nullptr, *j_t = nullptr, *j_b = nullptr;
};

namespace msbg {
namespace pband {

// r = b - A x over the listed cells (r optional) and the FP64 norms into
// p->d_norm, the refinement scale into p->d_scale.
void residual(msbg_poisson *p, hipStream_t s, int storage, const void *x, const void *b, float *r);
// Copies p->d_norm to the host (synchronizes the library stream).
int read_norms(msbg_poisson *p, double *r2, double *b2);

}  // namespace pband

namespace learned {

// One learned correction of x (FP32): network input from the residual in
// p->d_r (of x, with its norms in p->d_norm / p->d_scale) and the phase field,
// delta from the network, x_try = x + alpha delta with alpha minimising
// ||r - alpha A delta||, accepted when ||b - A x_try|| < ||b - A x||.  On
// return p->d_r / d_norm / d_scale hold the residual of the (possibly
// updated) x; r2 / b2 receive its norms.
int correct(msbg_poisson *p, msbg_learned *net, const msbg_grid *phi, float *x, const float *b, double r2_in,
            double *r2, double *b2, int *accepted);

}  // namespace learned
}  // namespace msbg
