// Model tables of the 2G-HDNNP path: host loader and the flat, POD layout
// the HIP kernels read.  The loader follows the setup sequence of
// Prediction::setup (Prediction.cpp:36-48): Mode::setupGeneric
// (Mode.cpp:213-237), setupSymmetryFunctionScaling (Mode.cpp:781-909) and
// setupNeuralNetworkWeights (Mode.cpp:1514-1547, 2331-2350).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nnpx {

constexpr int kMaxElem = 8;      // elements per model
constexpr int kMaxSF = 512;      // symmetry functions per element
constexpr int kMaxGroups = 64;   // SF groups per element
constexpr int kMaxMembers = 64;  // members per compact SF group
constexpr int kMaxMembersExp = 32;  // members per Gaussian (type 3/9) angular group
constexpr int kMaxLayers = 6;    // NN layers incl. input and output
constexpr int kMaxWidth = 128;   // NN layer width

enum ScalingType { ST_NONE = 0, ST_SCALE, ST_CENTER, ST_SCALECENTER, ST_SCALESIGMA };
enum Activation {
    AF_IDENTITY = 0, AF_TANH, AF_LOGISTIC, AF_SOFTPLUS, AF_RELU, AF_GAUSSIAN,
    AF_COS, AF_REVLOGISTIC, AF_EXP, AF_HARMONIC
};

// SF types: Gaussian radial (2), narrow (3) and wide (9) angular with a
// cutoff function; compact radial (20), narrow (21) and wide (22) angular.
__host__ __device__ inline bool sf_compact(int type) { return type >= 20; }
__host__ __device__ inline bool sf_radial(int type) { return type == 2 || type == 12 || type == 20 || type == 23; }
// Element-weighted types (12, 13, 23, 24, 25): every neighbour element
// contributes, weighted by its atomic number; e1 = e2 = -1 in their groups.
__host__ __device__ inline bool sf_weighted(int type) { return type == 12 || type == 13 || type >= 23; }

// One SF group: members share (central element, type, rc, e1[, e2]); compact
// groups share (central element, type, e1[, e2]) and span (rmin, rc = rmax).
struct GroupT {
    int type;        // 2, 3, 9, 20, 21, 22
    int e1, e2;      // neighbour elements (e1 <= e2)
    int nmem;        // members
    int mem0;        // first member in ModelT::mem
    int pad;
    double rc, rc2;
    double rcinv, rci, iw;  // cutoff constants (CutoffFunction.cpp:95-112)
    double rmin;     // compact: lower radial bound of the group
};

// One member (one symmetry function) of a group.
struct MemberT {
    int sf;          // SF index of the central element (row of G)
    int slot1;       // derivative slot for neighbour element e1
    int slot2;       // derivative slot for neighbour element e2 (angular)
    int slotE[kMaxElem];  // derivative slot per neighbour element (weighted types)
    int calcExp;     // recompute exp (eta or rs differ from previous member)
    int useIntPow;
    int zetaInt;
    int pad0, pad1;
    double eta, rs, lambda, zeta;
    double factorNorm;   // angular: 2^(1-zeta) * scaling factor
    double factorDeriv;  // angular: 2 eta / zeta / lambda
    double zetaLambda;   // angular: zeta * lambda
    double sfac;         // scaling factor (SymFnc::scalingFactor)
    double gnorm;        // angular: factorNorm / sfac
    // compact members (CompactFunction/CoreFunction): radial function on
    // (rl, rcm), angular function on (al, ar) radians
    int core;            // 0 exponential, 1..4 polynomial order
    int asym;            // asymmetric polynomial (x -> (2 - x) x)
    int calcComp;        // narrow: r_jk factor evaluated by this member
    int walk;            // narrow: the reused r_jk factor may come from another subtype
    double rl, rcm;      // radial window
    double cen, scl;     // radial centre and 1/half-width
    double al, ar;       // angular window (radians)
    double acen, ascl;   // angular centre and 1/half-width
};

// SymFnc::scale (SymFnc.cpp:169-195): Smin + factor * (v - shift) or v - shift.
struct ScaleT {
    int type;
    int pad;
    double Smin, Gmin, Gmean, factor;
    double Gmax;     // scaling.data maximum (extrapolation warnings)
};

struct NNT {
    int nl;                 // layers incl. input/output
    int normalize;          // normalize_nodes
    int n[kMaxLayers];      // neurons per layer
    int act[kMaxLayers];    // activation per layer
    int woff[kMaxLayers];   // offset of W_l (row-major [n_l][n_{l-1}]) in the weight buffer
    int boff[kMaxLayers];   // offset of b_l
};

struct ElementT {
    int nsf;               // symmetry functions
    int ngrp;              // groups (radial first, then angular)
    int grp0;              // first group in ModelT::grp
    int nrad;              // radial groups
    int sf0;               // first SF in ModelT::scale
    int ucut;              // 1: every group of the element has the same cutoff radius
    int nsfd[kMaxElem];    // derivative slots per neighbour element
    int tab0[kMaxElem];    // first entry of the slot->SF table per neighbour element
    NNT nn;
};

struct ModelT {
    int nelem;
    int cutoffType;
    int ngroups;
    int nmembers;
    int max_sf;             // max SF count over elements
    int max_sfd;            // max derivative slots (over central and neighbour element)
    int max_mem_ang;        // max members of a Gaussian angular group
    int max_mem_rad;
    int max_mem_comp;       // max members of a compact angular group
    int has_exp;            // any Gaussian (type 2/3/9) group
    int has_comp;           // any compact (type 20-25) group
    int has_weighted;       // any element-weighted (type 12/13/23/24/25) group
    double Zel[kMaxElem];   // atomic number by element index
    int max_width;          // max NN layer width (input layer included)
    int max_hidden;         // max width of the hidden and output layers
    int nweights;           // FP64 weights+biases in the weight buffer
    double maxCutoff;
    double alpha;
    double tanhPre;         // CT_TANH prefactor (CutoffFunction.cpp:28)
    ElementT el[kMaxElem];
    GroupT grp[kMaxElem * kMaxGroups];
    MemberT mem[kMaxElem * kMaxSF];
    ScaleT scale[kMaxElem * kMaxSF];
    int table[kMaxElem * kMaxElem * kMaxSF];  // slot -> SF index
};

// Host-side model: the device tables plus what the host needs.
struct HostModel {
    ModelT t{};
    std::vector<double> weights;        // NN parameters, FP64
    std::vector<std::string> symbols;   // by element index
    std::vector<int> Z;
    std::vector<double> energyOffset;
    bool normalize = false;
    double meanEnergy = 0.0, convEnergy = 1.0, convLength = 1.0;
};

// Parse input.nn, scaling.data and weights.%03d.data from dir.
// Throws nnpx::Error on failure.
void load_model(const std::string& dir, HostModel& m);

struct Error {
    int status;
    std::string msg;
};

}  // namespace nnpx
