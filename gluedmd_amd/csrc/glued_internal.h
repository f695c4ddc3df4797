// Internal types shared by the host engine and the HIP kernels of the GLUED
// enhanced-sampling step (CV evaluation -> bias evaluation -> chain-rule
// force scatter).  Nothing here crosses the C-ABI.
//
// HBM layout (all walker-major, one walker's block contiguous):
//   posq      float4[sum padded_w]       FP32 xyz + charge, OpenMM order
//   forces    int64 [3 * sum padded_w]   field-major per walker, 2^32 fixed point
//   gidx      int   [W][S]               GPU index of every referenced atom slot
//   cvVal     double[W][V]               CV values
//   jac{x,y,z} float[W][R]               Jacobian rows dCV/dx (row -> slot, cv)
//   bgrad     double[W][G]               dV_b/ds_d of every bias b, CV d
//   benergy   double[W][B]               bias energies
//   dUdCV     double[W][V]               chain-rule factors
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace glued {

// CV type codes: openmmapi/include/GluedForce.h:27-73.
enum CVType {
    CV_DISTANCE = 1,
    CV_ANGLE = 2,
    CV_DIHEDRAL = 3,
    CV_COM_DISTANCE = 4,
    CV_GYRATION = 5,
    CV_COORDINATION = 6,
    CV_RMSD = 7,
    CV_PATH = 10,
    CV_POSITION = 11,
    CV_DRMSD = 12,
    CV_CONTACTMAP = 13,
    CV_ENERGY = 23,
};

// Bias type codes: openmmapi/include/GluedForce.h:76-95, plus the hill-list
// metadynamics extension (no counterpart in the reference).
enum BiasType {
    BIAS_HARMONIC = 1,
    BIAS_METAD = 3,
    BIAS_PBMETAD = 4,
    BIAS_OPES = 5,
    BIAS_UPPER_WALL = 8,
    BIAS_LOWER_WALL = 9,
    BIAS_LINEAR = 10,
    BIAS_OPES_EXPANDED = 11,
    BIAS_OPES_MULTITHERMAL = 15,
    BIAS_METAD_HILLS = 101,
};

constexpr double FIXED_SCALE = 4294967296.0;  // 2^32, OpenMM's force fixed point
constexpr int MAX_BIAS_D = 8;                  // CVs per bias (reference: MAX_OPES_CVS 16)

struct WalkerInfo {
    long long posOff;    // float4 offset of this walker's posq
    long long forceOff;  // int64 offset of this walker's force block
    int padded;          // field stride of the force block
    int natoms;
};

// Common per-launch view of the walker-major state.
struct StepView {
    const float4* posq;
    long long* force;
    const WalkerInfo* walkers;
    const int* gidx;      // [W][S]
    const double* box;    // [W][12] = a, b, c, 1/ax, 1/by, 1/cz
    double* cvVal;        // [W][V]
    const int* rowPos;    // [R] storage position of plan row r (rows are stored grouped by atom slot)
    float4* jrow;         // [W][R] Jacobian rows: FP32 x, y, z (w unused), one 16-byte load or store per row
    double* bgrad;        // [W][G]
    double* benergy;      // [W][B]
    double* dUdCV;        // [W][V]
    const double* energyU; // [W] value of the CV_ENERGY input
    double* pathAux;       // [W][nPath][2] ds/dlambda, dz/dlambda
    double* pathMsd;       // [W][msdStride] per-frame MSD of every path
    double* pathPart;      // [W][partStride] MSD partials (cross terms, |x'|^2) per K slice
    const float* pathCentre;   // per path: the mean frame [3M] (FP32), the centre of the MSD contraction
    const double* pathR2;  // [msdStride] |r'_f|^2 of the centred frames (FP64)
    int nPath;
    int msdStride;
    int partStride;
    int gStride;          // entries per walker of the resolved group / path gather tables
    int pStride;
    int S, V, R, G, B;
    int pbc;
};

// Few-atom CVs (distance, angle, dihedral, position): one work item each.
struct SmallCVItem {
    int type;
    int slot[4];
    int value;
    int row0;
    int comp;  // position component
};

// Generic "group" CVs (gyration, COM distance, RMSD, DRMSD, contact map):
// atoms / pairs are listed in entry space (begin..end); one workgroup per
// (walker, item).  The gather table holds, per walker, the posq index of
// every entry (walker offset + GPU index), rebuilt on every atom reorder, so
// a gather is one dependent load (table) before the position load.
struct GroupCVItem {
    int type;
    int begin, end;       // range into the slot list / pair list
    int split;            // COM: size of group 1; RMSD: 1 = Kabsch fit
    int value;
    int row0;
    int refOff;           // RMSD reference (float3 units) / DRMSD ref / CM params
    double refCentroid[3];
    double refSum[3];     // Kabsch: sum of y - refCentroid over the reference (FP64, host)
    double refSq;         // Kabsch: sum of |y - refCentroid|^2
    int chunk0;           // contact map in chunks: first chunk (over all chunked maps)
    int chunkItem;        // contact map in chunks: index of its partial counter
};

// Contact maps longer than this many pairs are evaluated in chunks of this
// many pairs, one 256-thread workgroup per chunk (2 pairs per thread).
constexpr int CM_CHUNK = 512;

// Coordination CV with optional cutoff (cell list when the box allows it).
struct CoordItem {
    int aBegin, nA;       // slots of group A in coordSlots
    int bBegin, nB;       // slots of group B
    float r0, rn, rm;
    int intPow;           // 1 when n, m are small integers (fast integer powers)
    double rc2;           // cutoff^2 (0 = no cutoff)
    // switching constants formed once on the host (FP64, unfused):
    // 1/r0 and the Taylor coefficients used for |r/r0 - 1| < 1e-4
    double ir0, preRes, preDfunc, preSecDev;
    int cellBlock;        // threads per cell workgroup (256, 512, 1024: a tuned launch shape)
    double ir0sq;         // 1/r0^2
    int rat612;           // n = 6, m = 12: the switch is exactly 1/(1 + x^6)
    int value;
    int rowA, rowB;
    int useCells;         // cell list or brute force
    int nc[3];            // cells per dimension
    int cellBase;         // offset into the cell arrays
    int sortBase;         // offset into the sorted-atom arrays (A then B)
    int blkBase;          // offset into the per-workgroup cell-count table of the sort
};

// Path CV (s, z), MSD metric.
struct PathItem {
    int begin, M;         // slots
    int N;                // frames
    double lambda;        // FP64 (differentiable parameter)
    int frameOff;         // float offset of frame 0 (N*M*3 floats)
    int valueS;           // z at valueS + 1
    int row0;             // rows interleaved s/z per atom
    int aux;              // index of this path in the [W][nPath][2] d(s,z)/d(lambda) buffer
    int msdOff;           // offset of frame 0 in a walker's msd row ([W][msdStride] FP64)
    int msdBlk0;          // first unit of this path in the per-kind MSD launch
    int msdBlkF0;         // first block of this path in the MSD range of the fused CV launch
    int msdBlkF;          // blocks of this path per walker row of the fused CV launch
    int rowBlk0;          // first row block of this path (PATH_TILE atoms per block)
    int frameTiles;       // ceil(N / PATH_TF)
    int slices;           // K slices of the cross term (each PATH_KC * chunksPerSlice wide)
    int chunksPerSlice;
    int units;            // frameTiles * slices: MSD blocks per walker tile
    int partOff;          // offset in a walker's partial row: [slices][N] cross terms, then [slices] |x'|^2
    int centreOff;        // float offset of this path's centre (3M)
};

// Path CV work decomposition (shared by the host planner and the kernels).
// MSD: the cross term sum_k x'_wk r'_fk of the centred coordinates is a
// matrix-core contraction (v_mfma_f64_16x16x4f64) over tiles of PATH_TW
// walkers x PATH_TF frames, K in chunks of PATH_KC values (PATH_KC / 3 atoms).
constexpr int PATH_TW = 16;            // walkers per MSD tile (MFMA rows)
constexpr int PATH_TF = 16;            // frames per MSD tile (MFMA columns)
constexpr int PATH_KC = 96;            // K per chunk: 32 atoms x 3 components
constexpr int PATH_MAX_SLICES = 16;    // K slices (partials per frame) of one path
constexpr int PATH_TILE = 64;          // atoms per row block (one per lane)
constexpr int PATH_PHASES = 8;         // waves per row block, frames f = phase mod 8
constexpr int PATH_MAX_FRAMES = 4096;  // frames per path (LDS weight table)

}  // namespace glued

namespace glued {

// Grid metadynamics (BIAS_METAD).  CommonBiasKernels.cpp:206-297.
struct GridBiasDev {
    int D, cv[3], gslot, eslot;
    double origin[3], invSpacing[3], sigma[3];
    int npts[3], strides[3], periodic[3];
    int total;
    double h0, hif;           // heightInvFactor = 1/((gamma-1) kT), 0 if gamma <= 1
    double* grid;             // [W or 1][total]
    long long walkerStride;   // 0 when shared by all walkers
};

// Hill-list metadynamics (BIAS_METAD_HILLS, extension).
struct HillBiasDev {
    int D, cv[MAX_BIAS_D], gslot, eslot;
    double sigma[MAX_BIAS_D], period[MAX_BIAS_D], h0, hif;
    // host-computed reciprocals (the oracle forms the same expressions):
    // i2s2 = 1/(2 sigma^2), is2 = 1/sigma^2, invPeriod = 1/period
    double i2s2[MAX_BIAS_D], is2[MAX_BIAS_D], invPeriod[MAX_BIAS_D];
    float* centers;           // [W or 1][D][cap]  (SoA, coalesced per CV; FP32, widened before use)
    float* heights;           // [W or 1][cap]     (FP32)
    int* count;               // [W or 1]
    int cap;
    int shared;
    int chunks;               // blocks per walker in the eval launch
    double* part;             // [W][chunks][D+1]
    unsigned int* counter;    // [W]
    double* depPart;          // [T][workgroups][OPES_PART]: deposit partials (private tables)
    double* depRec;           // [T][record]: the deposit's change, for the sum_uprob launch
    unsigned int* depCount;   // [T][2]: deposit tickets
};

// OPES (BIAS_OPES).  CommonBiasKernels.cpp:489-664.
// Private OPES tables deposit over OPES_DEP_G workgroups in two launches:
// per workgroup a partial of OPES_PART doubles (best n2, its index, that
// kernel's centre, width and log-weight), and a record of OPES_REC doubles
// per table carries the change between the launches.
constexpr int OPES_DEP_G = 16;
constexpr int OPES_PART = 3 + 2 * MAX_BIAS_D;
constexpr int OPES_REC = 6 + 6 * MAX_BIAS_D;

struct OpesBiasDev {
    int D, cv[MAX_BIAS_D], gslot, eslot;
    int variant, adaptiveStride, cap, shared, chunks;
    double kT, gamma, invGF, cutoff2, eps, barrier, sigmaMin;
    double sigma0[MAX_BIAS_D];
    double* centers;          // [T][D][cap]   T = 1 (shared) or W tables
    double* sigmas;           // [T][D][cap]
    double* logw;             // [T][cap]
    int* ints;                // [T][4] = numKernels, numAllocated, stepCount, nSamples
    double* dbl;              // [T][4 + 2D] = sum_uprob, sumWeights, sum_w, sum_w2, mean[D], M2[D]
    double* part;             // [W][chunks][D+1]
    unsigned int* counter;    // [W]
    double* depPart;          // [T][workgroups][OPES_PART]: deposit partials (private tables)
    double* depRec;           // [T][record]: the deposit's change, for the sum_uprob launch
    unsigned int* depCount;   // [T][2]: deposit tickets
};

// Few-parameter biases evaluated by one thread (grid MetaD eval, restraints,
// OPES expanded, multithermal).
struct SmallBiasItem {
    int type, D, gslot, eslot;
    int cv[MAX_BIAS_D];
    int bias;                 // index into the grid-bias table (METAD); restraints: flag
                              // (HARMONIC: periodic; walls: 1 = lower); PBMETAD: first grid value
    int nState;               // OPES_EXPANDED / MULTITHERMAL states; PBMETAD: grid values per table
    int stateOff;             // offset into the state arrays (restraints, PBMETAD: their constants)
    double p[8];              // type-specific scalars (k, at, kT0, invKT, ...)
    GridBiasDev grid;         // METAD: a copy of the grid descriptor (one load less before the CV reads)
};

struct SmallBiasState {
    const GridBiasDev* grids;
    const double* stateConst; // logW (expanded) / beta - beta0 (multithermal)
    double* stateVar;         // [W or 1][...] deltaF (multithermal)
    double* scal;             // [W or 1][nSmall][4]: logZ, numUpdates, rct, counter
    long long stateVarStride, scalStride;
    double* pbGrid;           // PBMETAD sub-grids: [item base + table * values per table + sub-grid offset]
};

}  // namespace glued

namespace glued {

// Coordination scratch (cell lists, sorted atoms, per-block partial sums).
struct CoordBuffers {
    const int* slots;      // coordSlots
    const int* pos;        // [W][posStride] posq index of every coordSlots entry (walker offset included)
    int* cellOf;           // [W][cellOfStride]
    int* cellCount;        // [W][cellStride]
    int* cellStart;        // [W][cellStride + items]  (scan output, 1 extra per item)
    float4* sorted;        // [W][sortStride]  atoms in cell order (x, y, z, local index)
    float4* unsorted;      // [W][sortStride]  atoms in list order (staging of the cell build)
    int* rank;             // [W][sortStride]  rank of an atom among its sort workgroup's atoms of its cell
    int* blockCnt;         // [W][blkStride]   per (cell, sort workgroup) counts, then offsets
    long long* cnAcc;               // [W][items] CN: high part, units of 2^-24 (pass A adds, pass B converts and clears)
    unsigned long long* cnAccLo;    // [W][items] CN: low part, units of 2^-64
    unsigned long long* pairAcc;    // [W][items] pairs inside the cutoff (same life cycle)
    long long* npairs;              // [W][items]
    unsigned int* ticket;           // [W][items] group-A cell workgroups done with their CN sums
    int cellOfStride, cellStride, sortStride, blkStride, nItems, posStride;
};

// Wide group CVs launched as extra workgroups of another launch (the
// coordination rank kernel): items, atom slots, masses, references.
struct GroupExtra {
    const GroupCVItem* items;
    int n;
    const int* slots;
    const float* mass;
    const float* ref;
    // contact maps whose chunk partials the path-row launch adds up, in
    // its workgroup cmBlock of every walker row (-1: none)
    const GroupCVItem* cm;
    int nCm;
    int nCmBlocks;
    const double* cmPart;
    int cmBlock;
};

// Scatter plan: Jacobian rows grouped by atom slot (CSR), CV of every row,
// and the gradient slots that sum into each dU/dCV.
struct ScatterPlan {
    const int* segStart;     // [S+1] Jacobian rows of slot s are stored at [segStart[s], segStart[s+1])
    const int* entryCv;      // [R] CV value index of the row stored at each position
    const int* combStart;    // [V+1]
    const int* combSlot;     // [nComb]: bias-gradient slots summed into each CV's dU/dCV
    int nComb;
    const double* testGrad;  // [V] or null
    int energyValue;         // CV value index of the energy CV (-1 if none)
};

}  // namespace glued

namespace glued {

constexpr int MAX_CHUNKED = 16;

// The batched bias + scatter stage (one launch): chunk ranges of every
// chunked bias (hill lists first, then OPES tables), the one-thread biases,
// and the scatter plan for the fused scatter.
struct FusedBiasPlan {
    const HillBiasDev* hills;
    const OpesBiasDev* opes;
    int nHill, nOpes, nChunked;
    int chunkStart[MAX_CHUNKED + 1];
    int totalChunks;
    double* part;            // [W][totalChunks][MAX_BIAS_D + 1]
    unsigned int* counter;   // [W]
    const SmallBiasItem* small;
    int nSmall;
    SmallBiasState st;
    ScatterPlan sp;
    int scatter;             // 1: the finalizing workgroup also scatters
    int kind;                // chunk-kernel specialisation: 1 hills only, 2 OPES only (of one D), 0 mixed
    int kindD;               // that D
    // a copy of hills[0] passed by value: with one hill bias the chunk
    // workgroups read its fields from the kernel arguments, all in the first
    // round trip, instead of through a chain of loads from the table
    HillBiasDev hill0;
    // walkers per chunk workgroup (1, or CHUNK_WB when the only chunked bias
    // is one shared hill list: the chunk is loaded once for them all)
    int walkersPerChunk;
    int nWalkers;            // set at launch
    int extra;               // 1: some small bias is a restraint or PBMetaD (kernels with that code)
    int finalizeScatter;     // 1: the finalize launch also performs the slot scatter (a workgroup
                             // per BIAS_BLOCK slots, each forming dU/dCV itself)
};

constexpr int CHUNK_WB = 4;

}  // namespace glued

namespace glued {

// The batched CV stage (one launch for the group, path and few-atom CVs).
struct CVFusedPlan {
    const SmallCVItem* small;
    int nSmall;
    const GroupCVItem* group;
    int nGroup;
    const int* gslots;
    const float* gmass;
    const float* gref;
    const PathItem* path;
    int nPath;
    int nPathMsdBlocks;
    const int* pslots;
    const float* frames;
    const GroupCVItem* cm;  // contact maps in chunks
    int nCm;
    int nCmBlocks;          // chunks over all of them
    double* cmPart;         // [W][nCmBlocks] chunk partial sums
    unsigned int* cmCount;  // [W][nCm] chunks finished (the last resets it)
    int cmDefer;            // 1: the chunks only write their partials; the
                            // path-row launch adds them (no ticket, no fence)
};

}  // namespace glued
