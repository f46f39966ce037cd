// Training-gradient kernels: residuals, SF tangents along the force
// residual field, and the per-element network gradient.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace nnpx {


struct GradArgs {
    const double* w;       // FP64 weights (model layout), transposed copy after nweights
    int nweights;
    const double* G;       // [N*ldG] scaled SFs
    const double* Gdot;    // [N*ldG] tangents
    int ldG;
    const double* cE;      // [nstruct] 2 (E_s - E_ref,s) / N_s
    const int* sid;        // [N] structure of each atom
    const int* atoms;      // atoms grouped by element
    int blk_off[kMaxElem + 1];
    int atom_off[kMaxElem + 1];
    int nelem;
    int nconn[kMaxElem];   // connections per element
    int coff[kMaxElem];    // offset of each element's connections in grad
    int nconn_max;
    int maxw;              // max layer width
    int per_block;         // atoms of one element per block
    double* grad;          // [sum nconn], accumulated
    int ldgrad;            // 0: one sum; else grad + sid * ldgrad per structure
};

void launch_residual(const double* F, const double* Fref, const double* E, const double* Eref, const int* atom_off,
                     int natoms, int nstruct, double wF, double* u, double* cE, double* loss, hipStream_t stream);
void launch_tangent64(const ModelT* m, const SysView& s, const double* dGself, const double* dGnbr, int S,
                      const int* islot, const double* u, double* Gdot, int ldG, hipStream_t stream);
size_t grad_lds_bytes(const GradArgs& a, bool fp32);  // connection gradient + activations
size_t grad_act_bytes(const GradArgs& a, bool fp32);  // activations only (global-atomic variant)
void launch_grad(const ModelT* m, const GradArgs& a, int blocks, bool fp32, hipStream_t stream);

}  // namespace nnpx
