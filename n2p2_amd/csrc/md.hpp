// Velocity-Verlet integration kernels (device-side rebuild gate, per-step
// energies in a ring of kMDHistory entries).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace nnpx {

constexpr int kMDHistory = 1 << 16;  // per-step energies kept on the device

struct MDConst {
    double hk[kMaxElem];      // half-step velocity change per unit force, per element
    double half_m[kMaxElem];  // m/2 in units that give the kinetic energy in model energy units
    double dt;                // time step (positions advance by dt * v)
};

void md_kick_drift(double* pos, double* vel, const double* F, const int* elem, const MDConst& c, int n,
                   hipStream_t stream);
void md_check(const double* pos, const double* ref, int n, double lim2, int* gate, hipStream_t stream);
void md_wrap(double* pos, double* ref, const int* sid, const StructT* st, int n, const int* gate,
             hipStream_t stream);
void md_kick(double* vel, const double* F, const int* elem, const MDConst& c, int n, const int* counters,
             double* ekin, hipStream_t stream);
void md_step_end(int* gate, int* counters, const double* E, double* epot, double* ekin, hipStream_t stream);
void md_reset(int* gate, int* counters, double* ekin, hipStream_t stream);
void md_set_gate(int* gate, int v, hipStream_t stream);

}  // namespace nnpx
