// Host-side engine of the GLUED enhanced-sampling step: the force
// specification store (openmmapi/src/GluedForce.cpp), the plan builder and
// validation (CommonGluedKernels.cpp:3805-4577), the step (execute,
// :4646-5048; updateState, :5050-5157) and the bias-state blob
// (CommonSerialization.cpp).  Walker-major: one engine holds W walkers of
// the same force, each with its own positions, force block, CV values and
// private bias state; a bias may instead be shared by all walkers (the
// reference's multi-walker pointer sharing, CommonGluedKernels.cpp:5234-5262).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#include "glued_internal.h"

namespace glued {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);

// Device buffer with RAII ownership.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n)
    {
        release();
        if (n == 0) n = 8;
        hip_check(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
        hip_check(hipMemset(p, 0, n), "hipMemset");
    }
    template <class T>
    T* as() const
    {
        return static_cast<T*>(p);
    }
    template <class T>
    void upload(const std::vector<T>& v)
    {
        alloc(v.size() * sizeof(T));
        if (!v.empty()) hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
};

struct CVSpec {
    int type;
    std::vector<int> atoms;
    std::vector<double> params;
    int firstValue;
};

struct BiasSpec {
    int type;
    std::vector<int> cvs;
    std::vector<double> params;
    std::vector<int> iparams;
    int shared = 0;
};

struct ScatterPlanHost;

class Engine {
   public:
    Engine();
    ~Engine();

    // ---- specification (GluedForce.cpp:44-131) ----
    int addCV(int type, const int* atoms, int na, const double* params, int np);
    int addBias(int type, const int* cvs, int nc, const double* params, int np, const int* ip, int ni);
    void setShared(int bias, int shared);
    // one bias table shared by several handles: the device addresses of a
    // bias's table, and pointing this handle's table at another's
    std::vector<long long> multiWalkerPtrs(int bias);
    void redirectBias(int bias, const long long* ptrs, int n);
    void setTestBiasGradients(const double* g, int n);
    int numValues() const { return numValues_; }
    bool pbc = false;
    double temperature = -1.0;

    // ---- plan ----
    void initialize(int W, const int* natoms, const int* padded, int nMass, const double* masses);
    void setAtomOrder(int w, const int* gpuToUser, int n);
    void setBox(int w, const double* box9);
    void setEnergyInput(int w, double U, const float* Fext, hipStream_t s);

    // ---- step ----
    double execute(hipStream_t s, int w, const float4* posq, long long* force, bool forces, bool energy);
    void executeBatch(hipStream_t s, const float4* posq, long long* force, bool forces, bool useGraph);
    void executePlainAll(hipStream_t s, const float4* posq, long long* force, bool forces);
    void updateState(hipStream_t s, int step);
    // One MD step for every walker: updateState(step) then the batched
    // evaluation, replayed as one hipGraph per deposit pattern.
    void stepBatch(hipStream_t s, int step, const float4* posq, long long* force, bool forces);
    void executeStage(hipStream_t s, const float4* posq, long long* force, int stage);

    // ---- autograd support ----
    void copyBuffer(int which, void* dst, size_t bytes, hipStream_t s);
    void setHillHeightsDevice(int w, int bias, const double* d, int n, hipStream_t s);
    void setPathLambda(int path, double lambda);
    void setHillSigma(int bias, const double* sigma, int D);
    // part / partCap: a caller-owned buffer of per-block sigma partials
    // (doubles); nullptr = the engine's own, allocated on first use.
    void hillAdjoint(hipStream_t s, int bias, const double* gV, double* gH, double* gSigma, double* part = nullptr,
                     size_t partCap = 0);
    // doubles of per-block partials hillAdjoint needs for this bias
    size_t hillAdjointPartials(int bias) const;
    void setPathFrames(int path, const double* frames, int n);
    void opesAdjoint(hipStream_t s, int bias, const double* gV, double* gC, double* gS, double* gL, double* gU);
    void setOpesKernels(int w, int bias, const double* centers, const double* sigmas, const double* logw, int n,
                        double sumUprob);
    void pathFramesAdjoint(hipStream_t s, int path, const double* gV, double* gF);

    // ---- read-backs ----
    std::vector<double> cvValues(int w);
    std::vector<double> biasEnergies(int w);
    std::vector<double> dUdCV(int w);
    long long coordinationPairs(int w, int coordIdx);
    std::vector<char> getState(int w);
    void setState(int w, const std::vector<char>& blob);
    std::vector<double> opesMetrics(int w, int opesIdx);
    std::vector<float> kernelSigmas(int w, int opesIdx);
    void setHills(int w, int bias, const double* centers, const double* heights, int n);
    int getHills(int w, int bias, double* centers, double* heights, int cap);
    void setGrid(int w, int bias, const double* grid, int n);
    // Launch shapes picked by the autotuner (path MSD blocks per batch, path
    // row phases, coordination cell workgroup size) and how they were picked.
    struct Tuning {
        int pathBlocks = 512;  // target MSD blocks per path over the batch
        int pathPhases = 0;    // 4 or 8 waves per path row block (0 until planned)
        int coordBlock = 512;  // threads per coordination cell workgroup
    };
    std::string tuningReport() const;
    int numWalkers() const { return W_; }
    int jacobianRows() const { return R_; }
    int atomSlots() const { return S_; }
    double lastStepMs = 0.0;

   private:
    void validate(int minAtoms) const;
    Tuning tune_;
    bool tuned_ = false, evaluated_ = false;
    std::string tuneHow_ = "default", tuneLog_;
    void planPathWork();
    void applyTuning();
    void autotune(hipStream_t s);
    void adoptCachedTuning();
    std::string shapeKey() const;
    void buildPlan(int nMass, const double* masses);
    void launchStep(hipStream_t s, int w0, int nw, bool forces, int stageMask = 7, bool fused = false);
    StepView view(const float4* posq, long long* force) const;
    ScatterPlan scatterPlan() const;
    bool tinyPlan() const;
    void uploadGidx(int w);
    void sync(hipStream_t s);

    std::vector<CVSpec> cvs_;
    std::vector<BiasSpec> biases_;
    std::vector<double> testGrads_;
    int numValues_ = 0;
    bool initialized_ = false;

    // plan sizes
    int W_ = 0, S_ = 0, V_ = 0, R_ = 0, G_ = 0, B_ = 0;
    std::vector<int> natoms_, padded_;
    std::vector<long long> posOff_, forceOff_;
    std::vector<int> slotAtom_;          // user atom of every slot
    int groupBigNT_ = 1024;  // threads per workgroup of the wide group CVs (fixed per plan)
    std::vector<int> groupSlots_, pathSlots_, coordSlots_;  // slot of every group / path / coordination gather entry
    std::vector<std::vector<int>> userToGpu_;
    std::vector<double> boxHost_;        // [W][12]
    std::vector<int> rowSlot_, rowCv_;

    // items
    std::vector<SmallCVItem> small_;
    std::vector<GroupCVItem> group_, groupBig_;  // groupBig_: more than 1,024 atoms / pairs
    std::vector<GroupCVItem> cm_;                // contact maps of more than CM_CHUNK pairs, in chunks
    int cmBlocks_ = 0;
    std::vector<CoordItem> coord_;
    std::vector<PathItem> path_;
    int energyValue_ = -1;
    std::vector<double> energyU_;
    std::vector<const float*> energyF_;

    // bias plan
    struct GridHost {
        GridBiasDev d;
        int pace, spec;
        std::vector<int> numDeposited;  // per table
        DBuf grid;
    };
    struct HillHost {
        HillBiasDev d;
        int pace, spec;
        DBuf centers, heights, count, part, counter;
    };
    struct OpesHost {
        OpesBiasDev d;
        int pace, spec;
        std::vector<int> depositMirror;  // per table, the reference's CPU numKernels mirror
        std::vector<int> nSamplesMirror;
        DBuf centers, sigmas, logw, ints, dbl, part, counter, depPart, depRec, depCount;
    };
    struct SmallHost {
        int spec, pace, type;
    };
    // PBMetaD: N 1-D sub-grids per bias in dPbGrid_ (item base + table *
    // stride + sub-grid offset); the deposit count per table
    struct PbHost {
        int item, N, T, pace;
        std::vector<int> npts;
        long long base, stride;
        std::vector<int> numDeposited;
    };
    std::vector<PbHost> pbs_;
    long long pbGridTotal_ = 0;
    DBuf dPbGrid_;
    std::vector<GridHost*> grids_;
    std::vector<HillHost*> hills_;
    std::vector<OpesHost*> opes_;
    void bumpWelfordMirrors(OpesHost* o, int w0, int nw);
    // Deposit sites of updateState (bit per OPES table, OPES-expanded /
    // multithermal item, grid, hill list, in that order): which deposit at
    // `step`; the launches for a mask; the host mirrors for a mask.
    int depositSites() const;
    uint64_t depositMask(int step) const;
    void launchDeposits(hipStream_t s, uint64_t mask);
    void commitDeposits(uint64_t mask);
    void dropGraphs();
    void layoutCellTables();
    bool capturing_ = false;
    std::vector<SmallBiasItem> smallBias_;
    std::vector<SmallHost> smallBiasHost_;
    std::vector<double> smallConst_;
    int smallVarPerWalker_ = 0;
    std::vector<int> combStart_, combSlot_;

    // device buffers
    DBuf dWalkers_, dGidx_, dBox_, dCv_, dJ_, dBgrad_, dBener_, dUd_;
    DBuf dSmall_, dGroup_, dGroupBig_, dGroupPos_, dGroupMass_, dGroupRef_;
    DBuf dCm_, dCmPart_, dCmCount_;
    DBuf dCoord_, dCoordSlots_, dCoordPos_, dCellOf_, dCellCount_, dCellStart_, dSorted_, dUnsorted_, dCnAcc_, dCnAccLo_, dPairAcc_,
        dNpairs_, dRank_, dBlockCnt_, dCoordTicket_;
    int cellOfStride_ = 0, cellStride_ = 0, sortStride_ = 0, blkStride_ = 0;
    DBuf dPath_, dPathPos_, dPathFrames_, dPathW_;
    DBuf dSmallBias_, dGridTable_, dSmallConst_, dSmallVar_, dSmallScal_;
    DBuf dSegStart_, dSegRow_ /* rowPos */, dRowCv_ /* entryCv */, dCombStart_, dCombSlot_, dTestGrad_, dEnergyU_, dEnergyFPtr_;
    DBuf dHillTable_, dOpesTable_, dFusedPart_, dFusedCounter_, dPathAux_, dAdjPart_, dPathMsd_, dPathCoef_;
    int msdStride_ = 0, pathMsdBlocks_ = 0, pathRowBlocks_ = 0;
    int pathMsdUnits_ = 0, partStride_ = 0;
    // per path: centre [3M] (FP32) and |r_f - c|^2 (FP64) of the MSD contraction
    std::vector<float> pathCentreHost_;
    std::vector<double> pathR2Host_;
    DBuf dPathPart_, dPathCentre_, dPathR2_;
    void centrePath(int pi, const float* frames);
    int pathMaxN_ = 0;  // most frames of any path
    FusedBiasPlan fplan_{};

    // graph cache for the batched step
    hipGraphExec_t graphExec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    const void* graphPos_ = nullptr;
    const void* graphForce_ = nullptr;
    hipStream_t captureStream_ = nullptr;
    bool graphForces_ = true;
    // step graphs (deposits + evaluation), keyed by deposit mask
    std::map<uint64_t, std::pair<hipGraph_t, hipGraphExec_t>> stepGraphs_;
    const void* stepGraphPos_ = nullptr;
    const void* stepGraphForce_ = nullptr;
    bool stepGraphForces_ = true;

    bool cvReady_ = false;
    int lastUpdateStep_ = -1;
    int maxPadded_ = 0;
    // combined dense + slot scatter: per walker, {bit per GPU atom with rows,
    // rows-atoms in earlier words} per 32-atom word, and the slots of those
    // atoms in GPU order
    DBuf dSlotMask_, dSlotList_;   // [W][maxPadded/32] uint2, [W][S] int2 row range
    std::vector<int> segStart_;    // [S + 1] row range of every slot (host copy)
    int maskStride_ = 0;
    const float4* lastPos_ = nullptr;
    long long* lastForce_ = nullptr;
};

}  // namespace glued
