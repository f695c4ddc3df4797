// Host engine of the GLUED enhanced-sampling step (see engine.hpp).
#include "engine.hpp"
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

namespace glued {

// kernel launchers (kernels_*.hip)
void launch_cv_small(const StepView&, const SmallCVItem*, int, int, int, hipStream_t);
void launch_cv_group(const StepView&, const GroupCVItem*, int, const int*, const float*, const float*, int, int,
                     hipStream_t);
void launch_cv_path(const StepView&, const PathItem*, int, int, int, int, const int*, const float*, int, int, int,
                    hipStream_t);
void launch_cv_group_big(const StepView&, const GroupCVItem*, int, const int*, const float*, const float*, int, int,
                         int, hipStream_t);
bool launch_path_rows(const StepView&, const PathItem*, int, int, int, const int*, const float*, int, int, int,
                      hipStream_t, const GroupExtra*);

bool launch_coordination(const StepView&, const CoordBuffers&, const CoordItem*, const CoordItem&, int, int, int,
                         hipStream_t, const GroupExtra*);
void launch_bias_small(const StepView&, const SmallBiasState&, const SmallBiasItem*, int, int, int, hipStream_t);
void launch_bias_hills(const StepView&, const HillBiasDev&, int, int, hipStream_t);
void launch_bias_opes(const StepView&, const OpesBiasDev&, int, int, hipStream_t);
void launch_dep_grid(const StepView&, const GridBiasDev&, int, int, hipStream_t);
void launch_dep_pb(const StepView&, const SmallBiasState&, const SmallBiasItem&, int, int, hipStream_t);
void launch_dep_hills(const StepView&, const HillBiasDev&, int, int, hipStream_t);
void launch_f64_to_f32(const double*, float*, int, hipStream_t);
void launch_opes_adjoint(const StepView&, const OpesBiasDev&, const double*, double*, double*, double*, double*, int,
                         hipStream_t);
void launch_dep_opes(const StepView&, const OpesBiasDev&, int, int, hipStream_t);
void launch_opes_welford(const StepView&, const OpesBiasDev&, int, int, hipStream_t);
void launch_dep_small(const StepView&, const SmallBiasState&, const SmallBiasItem*, int, int, int, int, hipStream_t);
int hill_chunks_for(int);
int opes_chunks_for(int);

void launch_scatter(const StepView&, const ScatterPlan&, int, int, int, hipStream_t);
void launch_scatter_dense(const StepView&, const float* const*, int, int, int, int, hipStream_t);
void launch_scatter_dense_rows(const StepView&, const ScatterPlan&, const float* const*, int, const int2*,
                               const uint2*, int, int, int, int, int, hipStream_t);
void launch_step_tiny(const StepView&, const SmallCVItem*, int, const FusedBiasPlan&, int, hipStream_t);
void launch_step_bias(const StepView&, const FusedBiasPlan&, int, hipStream_t);
void launch_step_bias_chunks(const StepView&, const FusedBiasPlan&, int, hipStream_t);
void launch_cv_fused(const StepView&, const CVFusedPlan&, int, hipStream_t);
void launch_cv_cm_chunks(const StepView&, const CVFusedPlan&, int, int, hipStream_t);
void launch_hill_adjoint(const StepView&, const HillBiasDev&, const double*, double*, double*, double*, int,
                         hipStream_t);
int hill_adjoint_blocks(int);
void launch_path_frames_adjoint(const StepView&, const PathItem*, int, int, int, const int*, const float*,
                                const double*, double*, double*, int, hipStream_t);

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

static void fail(const std::string& m) { throw Error(m); }

Engine::Engine() {}

Engine::~Engine()
{
    dropGraphs();
    if (captureStream_) (void)hipStreamDestroy(captureStream_);
    for (auto* g : grids_) delete g;
    for (auto* h : hills_) delete h;
    for (auto* o : opes_) delete o;
}

// ---------------------------------------------------------------------------
// Specification: GluedForce.cpp:44-107.  PATH occupies two value slots.
// ---------------------------------------------------------------------------
int Engine::addCV(int type, const int* atoms, int na, const double* params, int np)
{
    if (initialized_) fail("addCollectiveVariable: force already initialized");
    CVSpec c;
    c.type = type;
    c.atoms.assign(atoms, atoms + na);
    c.params.assign(params, params + np);
    c.firstValue = numValues_;
    numValues_ += (type == CV_PATH) ? 2 : 1;
    cvs_.push_back(c);
    return c.firstValue;
}

int Engine::addBias(int type, const int* cvs, int nc, const double* params, int np, const int* ip, int ni)
{
    if (initialized_) fail("addBias: force already initialized");
    for (int i = 0; i < nc; i++)
        if (cvs[i] < 0 || cvs[i] >= numValues_)
            fail("addBias: CV index " + std::to_string(cvs[i]) + " out of range [0," + std::to_string(numValues_) +
                 ")");
    BiasSpec b;
    b.type = type;
    b.cvs.assign(cvs, cvs + nc);
    b.params.assign(params, params + np);
    b.iparams.assign(ip, ip + ni);
    biases_.push_back(b);
    return (int)biases_.size() - 1;
}

void Engine::setShared(int bias, int shared)
{
    if (initialized_) fail("share_bias: call before initialize");
    if (bias < 0 || bias >= (int)biases_.size()) fail("share_bias: bias index out of range");
    biases_[bias].shared = shared ? 1 : 0;
}

// Cross-handle sharing (CommonGluedKernels.cpp:5234-5262 and the kernel
// interface's getMultiWalkerPtrs, GluedKernels.h:89-104).  A table is shared
// whole: for OPES the kernels, the counts and the normalisation sums (sum_uprob,
// KDNorm, sample count), so one-walker handles pointed at one primary and
// stepped in turn on one stream give the bits of one handle whose walkers
// share the table.  Only single-table biases take part (one walker, or a bias
// shared by all walkers of the handle).
namespace {
void require_device_range(long long ptr, size_t bytes, const char* what)
{
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (ptr == 0 || hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)(uintptr_t)ptr) != hipSuccess) {
        (void)hipGetLastError();
        fail(std::string("redirect_bias: ") + what + " is not device memory of this process");
    }
    const uintptr_t lo = (uintptr_t)base, p = (uintptr_t)ptr;
    if (p < lo || p + bytes > lo + size)
        fail(std::string("redirect_bias: ") + what + " is smaller than this bias's table (different shape?)");
}
}  // namespace

std::vector<long long> Engine::multiWalkerPtrs(int bias)
{
    if (!initialized_) fail("multiwalker_ptrs: initialize first");
    if (bias < 0 || bias >= (int)biases_.size()) fail("multiwalker_ptrs: bias index out of range");
    const int T = biases_[bias].shared ? 1 : W_;
    if (T != 1) fail("multiwalker_ptrs: the bias has one table per walker; share it or use one walker");
    for (auto* g : grids_)
        if (g->spec == bias) return {(long long)(uintptr_t)g->d.grid};
    for (auto* h : hills_)
        if (h->spec == bias)
            return {(long long)(uintptr_t)h->d.centers, (long long)(uintptr_t)h->d.heights,
                    (long long)(uintptr_t)h->d.count};
    for (auto* o : opes_)
        if (o->spec == bias)
            return {(long long)(uintptr_t)o->d.centers, (long long)(uintptr_t)o->d.sigmas,
                    (long long)(uintptr_t)o->d.logw, (long long)(uintptr_t)o->d.ints,
                    (long long)(uintptr_t)o->d.dbl};
    fail("multiwalker_ptrs: only METAD, METAD_HILLS and OPES tables can be shared between handles");
}

void Engine::redirectBias(int bias, const long long* p, int n)
{
    if (!initialized_) fail("redirect_bias: initialize first");
    if (bias < 0 || bias >= (int)biases_.size()) fail("redirect_bias: bias index out of range");
    const int T = biases_[bias].shared ? 1 : W_;
    if (T != 1) fail("redirect_bias: the bias has one table per walker; share it or use one walker");
    auto want = [&](int k) {
        if (n != k) fail("redirect_bias: expected " + std::to_string(k) + " pointers, got " + std::to_string(n));
    };
    bool done = false;
    for (size_t k = 0; k < grids_.size() && !done; k++) {
        GridHost* g = grids_[k];
        if (g->spec != bias) continue;
        want(1);
        require_device_range(p[0], sizeof(double) * (size_t)g->d.total, "grid");
        g->d.grid = (double*)(uintptr_t)p[0];
        hip_check(hipMemcpy(dGridTable_.as<GridBiasDev>() + k, &g->d, sizeof(GridBiasDev), hipMemcpyHostToDevice),
                  "redirect_bias: grid table");
        for (size_t i = 0; i < smallBias_.size(); i++) {
            SmallBiasItem& it = smallBias_[i];
            if (it.type != BIAS_METAD || it.bias != (int)k) continue;
            it.grid = g->d;
            hip_check(hipMemcpy(dSmallBias_.as<SmallBiasItem>() + i, &it, sizeof(SmallBiasItem),
                                hipMemcpyHostToDevice),
                      "redirect_bias: small-bias table");
        }
        done = true;
    }
    for (auto* h : hills_) {
        if (done || h->spec != bias) continue;
        want(3);
        require_device_range(p[0], sizeof(float) * (size_t)h->d.D * h->d.cap, "hill centres");
        require_device_range(p[1], sizeof(float) * (size_t)h->d.cap, "hill heights");
        require_device_range(p[2], sizeof(int), "hill count");
        h->d.centers = (float*)(uintptr_t)p[0];
        h->d.heights = (float*)(uintptr_t)p[1];
        h->d.count = (int*)(uintptr_t)p[2];
        done = true;
    }
    for (auto* o : opes_) {
        if (done || o->spec != bias) continue;
        want(5);
        const size_t D = (size_t)o->d.D, cap = (size_t)o->d.cap;
        require_device_range(p[0], sizeof(double) * D * cap, "OPES centres");
        require_device_range(p[1], sizeof(double) * D * cap, "OPES widths");
        require_device_range(p[2], sizeof(double) * cap, "OPES log-weights");
        require_device_range(p[3], sizeof(int) * 4, "OPES counts");
        require_device_range(p[4], sizeof(double) * (4 + 2 * D), "OPES sums");
        o->d.centers = (double*)(uintptr_t)p[0];
        o->d.sigmas = (double*)(uintptr_t)p[1];
        o->d.logw = (double*)(uintptr_t)p[2];
        o->d.ints = (int*)(uintptr_t)p[3];
        o->d.dbl = (double*)(uintptr_t)p[4];
        done = true;
    }
    if (!done) fail("redirect_bias: only METAD, METAD_HILLS and OPES tables can be shared between handles");
    // the chunked-bias tables and the graphs hold the old addresses
    for (size_t k = 0; k < hills_.size(); k++)
        hip_check(hipMemcpy(dHillTable_.as<HillBiasDev>() + k, &hills_[k]->d, sizeof(HillBiasDev),
                            hipMemcpyHostToDevice),
                  "redirect_bias: hill table");
    if (!hills_.empty()) fplan_.hill0 = hills_[0]->d;
    for (size_t k = 0; k < opes_.size(); k++)
        hip_check(hipMemcpy(dOpesTable_.as<OpesBiasDev>() + k, &opes_[k]->d, sizeof(OpesBiasDev),
                            hipMemcpyHostToDevice),
                  "redirect_bias: OPES table");
    dropGraphs();
}

void Engine::setTestBiasGradients(const double* g, int n)
{
    if (initialized_) fail("setTestBiasGradients: must be called before initialization");
    testGrads_.assign(g, g + n);
}

// ---------------------------------------------------------------------------
// Validation: CommonGluedKernels.cpp:3879-3927 (atom ranges, minimum atom and
// parameter counts), CommonBiasKernels.cpp (bias parameter layouts).
// ---------------------------------------------------------------------------
void Engine::validate(int minAtoms) const
{
    for (size_t i = 0; i < cvs_.size(); i++) {
        const CVSpec& c = cvs_[i];
        auto err = [&](const std::string& m) { fail("GluedForce CV " + std::to_string(i) + ": " + m); };
        auto needA = [&](size_t n) {
            if (c.atoms.size() < n)
                err("expected >= " + std::to_string(n) + " atoms, got " + std::to_string(c.atoms.size()));
        };
        auto needP = [&](size_t n) {
            if (c.params.size() < n)
                err("expected >= " + std::to_string(n) + " parameters, got " + std::to_string(c.params.size()));
        };
        const bool coord = c.type == CV_COORDINATION;
        for (size_t k = coord ? 1 : 0; k < c.atoms.size(); k++)
            if (c.atoms[k] < 0 || c.atoms[k] >= minAtoms)
                err("atom index " + std::to_string(c.atoms[k]) + " out of range [0," + std::to_string(minAtoms) + ")");
        switch (c.type) {
            case CV_DISTANCE: needA(2); break;
            case CV_ANGLE: needA(3); break;
            case CV_DIHEDRAL: needA(4); break;
            case CV_COM_DISTANCE: needA(1); needP(1); break;
            case CV_GYRATION: needA(1); break;
            case CV_COORDINATION:
                needA(2);
                needP(3);
                if (c.atoms[0] < 1 || c.atoms[0] >= (int)c.atoms.size())
                    err("coordination group-A size atoms[0]=" + std::to_string(c.atoms[0]) + " invalid");
                break;
            case CV_RMSD: needA(1); needP(3 * c.atoms.size()); break;
            case CV_PATH: {
                needP(2);
                int N = (int)c.params[1];
                if (N < 2) err("path needs >= 2 frames");
                if (N > PATH_MAX_FRAMES)
                    err("path frame count " + std::to_string(N) + " exceeds " + std::to_string(PATH_MAX_FRAMES));
                needP(2 + (size_t)N * 3 * c.atoms.size());
                break;
            }
            case CV_POSITION: needA(1); needP(1); break;
            case CV_DRMSD: needA(2); needP(c.atoms.size() / 2); break;
            case CV_CONTACTMAP: needA(2); needP(4 * (c.atoms.size() / 2)); break;
            case CV_ENERGY: break;
            default: err("CV type " + std::to_string(c.type) + " is not supported by this implementation");
        }
    }
    for (size_t b = 0; b < biases_.size(); b++) {
        const BiasSpec& s = biases_[b];
        auto err = [&](const std::string& m) { fail("GluedForce bias " + std::to_string(b) + ": " + m); };
        const int D = (int)s.cvs.size();
        if (D < 1) err("needs at least one CV");
        switch (s.type) {
            case BIAS_METAD:
                if (D > 3) err("grid metadynamics supports at most 3 CVs");
                if ((int)s.params.size() < 3 + 3 * D) err("METAD params must be [h0, sigma*D, gamma, kT, min*D, max*D]");
                if ((int)s.iparams.size() < 1 + 2 * D) err("METAD int params must be [pace, bins*D, periodic*D]");
                break;
            case BIAS_METAD_HILLS:
                if (D > 4) err("hill metadynamics supports at most 4 CVs");
                if ((int)s.params.size() < 3 + 2 * D) err("hill METAD params must be [h0, sigma*D, gamma, kT, period*D]");
                if ((int)s.iparams.size() < 2) err("hill METAD int params must be [pace, maxHills]");
                break;
            case BIAS_OPES:
                if (D > 4) err("OPES supports at most 4 CVs");
                if ((int)s.params.size() < 3 + D) err("OPES params must be [kT, gamma, sigma0*D, sigmaMin]");
                // CommonBiasKernels.cpp:568-570
                if (!s.iparams.empty() && s.iparams[0] == 2 && (!std::isfinite(s.params[1]) || s.params[1] <= 1.0))
                    err("OPES_METAD_EXPLORE (variant=2) requires finite gamma > 1");
                break;
            case BIAS_OPES_EXPANDED:
                if ((int)s.params.size() < 1 + D) err("OPES_EXPANDED params must be [kT, w*D]");
                break;
            case BIAS_OPES_MULTITHERMAL: {
                if ((int)s.params.size() < 2)
                    err("BIAS_OPES_MULTITHERMAL needs at least one temperature (params = [kT0, beta_0, ...])");
                bool ok = false;
                for (auto& c : cvs_)
                    if (c.type == CV_ENERGY && c.firstValue == s.cvs[0]) ok = true;
                if (!ok) err("BIAS_OPES_MULTITHERMAL: cvIndices[0] must reference a CV_ENERGY value index");
                break;
            }
            case BIAS_HARMONIC:
                if (D > MAX_BIAS_D) err("harmonic restraint supports at most 8 CVs");
                if ((int)s.params.size() < 2 * D) err("HARMONIC params must be [k_0, s0_0, k_1, s0_1, ...]");
                break;
            case BIAS_LINEAR:
                if (D > MAX_BIAS_D) err("linear coupling supports at most 8 CVs");
                if ((int)s.params.size() < D) err("LINEAR params must be [k_0, k_1, ...]");
                break;
            case BIAS_PBMETAD:
                if (D > MAX_BIAS_D) err("PBMetaD supports at most 8 CVs");
                if ((int)s.params.size() < 3 + 3 * D) err("PBMETAD params must be [h0, gamma, kT, (sigma, min, max)*D]");
                if ((int)s.iparams.size() < 1 + 2 * D) err("PBMETAD int params must be [pace, (bins, periodic)*D]");
                for (int d = 0; d < D; d++) {
                    if (s.iparams[1 + 2 * d] < 1) err("PBMETAD needs at least one bin per CV");
                    if (!(s.params[5 + 3 * d] > s.params[4 + 3 * d])) err("PBMETAD grid max must exceed min");
                }
                if (s.iparams[0] < 1) err("PBMETAD pace must be positive");
                break;
            case BIAS_UPPER_WALL:
            case BIAS_LOWER_WALL:
                if (D > MAX_BIAS_D) err("walls support at most 8 CVs");
                if ((int)s.params.size() < 4 * D) err("wall params must be [at, kappa, eps, n] per CV");
                break;
            default: err("bias type " + std::to_string(s.type) + " is not supported by this implementation");
        }
    }
    // every bias owns at most one bit of the 64-bit deposit mask
    if (biases_.size() > 64) fail("GluedForce: at most 64 biases per force");
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
void Engine::initialize(int W, const int* natoms, const int* padded, int nMass, const double* masses)
{
    if (initialized_) fail("initialize: already initialized");
    if (W < 1) fail("initialize: need at least one walker");
    W_ = W;
    natoms_.assign(natoms, natoms + W);
    padded_.assign(padded, padded + W);
    int minAtoms = *std::min_element(natoms_.begin(), natoms_.end());
    for (int w = 0; w < W; w++)
        if (padded_[w] < natoms_[w]) fail("initialize: padded atom count smaller than atom count");
    validate(minAtoms);
    bool hasEnergy = false;
    for (const auto& cv : cvs_) hasEnergy |= cv.type == CV_ENERGY;
    if (hasEnergy)  // the dense energy-CV scatter moves 4 field entries per 16-byte access
        for (int w = 0; w < W; w++)
            if (padded_[w] % 4) fail("initialize: with an energy CV every padded atom count must be a multiple of 4");
    posOff_.resize(W);
    forceOff_.resize(W);
    long long po = 0;
    maxPadded_ = 0;
    for (int w = 0; w < W; w++) {
        posOff_[w] = po;
        forceOff_[w] = 3 * po;
        po += padded_[w];
        maxPadded_ = std::max(maxPadded_, padded_[w]);
    }
    buildPlan(nMass, masses);
    initialized_ = true;
}

void Engine::buildPlan(int nMass, const double* masses)
{
    // ---- atom slots ----
    std::map<int, int> slotOf;
    auto slot = [&](int atom) {
        auto it = slotOf.find(atom);
        if (it != slotOf.end()) return it->second;
        int s = (int)slotAtom_.size();
        slotAtom_.push_back(atom);
        slotOf[atom] = s;
        return s;
    };
    auto massOf = [&](int atom) -> float {
        if (masses && atom < nMass) return (float)masses[atom];
        return 1.0f;
    };
    V_ = numValues_;
    std::vector<int> groupSlots, coordSlots, pathSlots;
    std::vector<float> groupMass, groupRef, pathFrames;
    auto addRow = [&](int s, int cv) {
        rowSlot_.push_back(s);
        rowCv_.push_back(cv);
        return (int)rowSlot_.size() - 1;
    };
    for (const CVSpec& c : cvs_) {
        const int v = c.firstValue;
        switch (c.type) {
            case CV_DISTANCE:
            case CV_ANGLE:
            case CV_DIHEDRAL:
            case CV_POSITION: {
                SmallCVItem it{};
                it.type = c.type;
                it.value = v;
                int na = c.type == CV_DISTANCE ? 2 : c.type == CV_ANGLE ? 3 : c.type == CV_DIHEDRAL ? 4 : 1;
                it.row0 = (int)rowSlot_.size();
                for (int a = 0; a < na; a++) {
                    it.slot[a] = slot(c.atoms[a]);
                    addRow(it.slot[a], v);
                }
                it.comp = c.type == CV_POSITION ? (int)c.params[0] : 0;
                if (c.type == CV_POSITION && (it.comp < 0 || it.comp > 2)) fail("position component must be 0, 1 or 2");
                small_.push_back(it);
                break;
            }
            case CV_ENERGY: {
                if (energyValue_ >= 0) fail("only one CV_ENERGY is supported");
                energyValue_ = v;
                SmallCVItem it{};
                it.type = CV_ENERGY;
                it.value = v;
                it.row0 = -1;
                small_.push_back(it);
                break;
            }
            case CV_GYRATION:
            case CV_COM_DISTANCE:
            case CV_RMSD: {
                GroupCVItem it{};
                it.type = c.type;
                it.value = v;
                it.begin = (int)groupSlots.size();
                it.row0 = (int)rowSlot_.size();
                const int n = (int)c.atoms.size();
                for (int a = 0; a < n; a++) {
                    int s = slot(c.atoms[a]);
                    groupSlots.push_back(s);
                    groupMass.push_back(massOf(c.atoms[a]));
                    addRow(s, v);
                }
                it.end = (int)groupSlots.size();
                if (c.type == CV_COM_DISTANCE) {
                    it.split = (int)std::lround(c.params[0]);
                    if (it.split < 1 || it.split >= n) fail("COM distance: group-1 size out of range");
                }
                if (c.type == CV_RMSD) {
                    // params = reference xyz (float, as the reference stores it), optional
                    // trailing fit flag: 1 = optimal rotation + translation (Kabsch)
                    it.split = ((int)c.params.size() > 3 * n && c.params[3 * n] != 0.0) ? 1 : 0;
                    if ((groupRef.size() % 3) != 0) groupRef.resize((groupRef.size() / 3 + 1) * 3, 0.f);
                    it.refOff = (int)(groupRef.size() / 3);
                    double cs[3] = {0, 0, 0};
                    for (int k = 0; k < 3 * n; k++) {
                        float f = (float)c.params[k];
                        groupRef.push_back(f);
                        cs[k % 3] += (double)f;
                    }
                    for (int k = 0; k < 3; k++) it.refCentroid[k] = cs[k] / n;
                    // the reference moments of the Kabsch covariance do not change
                    // between evaluations: formed once here, in FP64
                    for (int k = 0; k < 3; k++) it.refSum[k] = 0.0;
                    it.refSq = 0.0;
                    for (int j = 0; j < n; j++)
                        for (int k = 0; k < 3; k++) {
                            const double y = (double)(float)c.params[3 * j + k] - it.refCentroid[k];
                            it.refSum[k] += y;
                            it.refSq += y * y;
                        }
                }
                group_.push_back(it);
                break;
            }
            case CV_DRMSD:
            case CV_CONTACTMAP: {
                GroupCVItem it{};
                it.type = c.type;
                it.value = v;
                it.begin = (int)groupSlots.size();
                it.row0 = (int)rowSlot_.size();
                const int np = (int)c.atoms.size() / 2;
                for (int p = 0; p < np; p++) {
                    int sa = slot(c.atoms[2 * p]), sb = slot(c.atoms[2 * p + 1]);
                    groupSlots.push_back(sa);
                    groupSlots.push_back(sb);
                    groupMass.push_back(1.f);
                    groupMass.push_back(1.f);
                    addRow(sa, v);
                    addRow(sb, v);
                }
                it.end = (int)groupSlots.size();
                it.refOff = (int)groupRef.size();
                const int per = c.type == CV_DRMSD ? 1 : 4;
                for (int k = 0; k < per * np; k++) groupRef.push_back((float)c.params[k]);
                group_.push_back(it);
                break;
            }
            case CV_COORDINATION: {
                CoordItem it{};
                const int nA = c.atoms[0];
                const int nB = (int)c.atoms.size() - 1 - nA;
                it.value = v;
                it.nA = nA;
                it.nB = nB;
                it.aBegin = (int)coordSlots.size();
                it.rowA = (int)rowSlot_.size();
                for (int a = 0; a < nA; a++) {
                    int s = slot(c.atoms[1 + a]);
                    coordSlots.push_back(s);
                    addRow(s, v);
                }
                it.bBegin = (int)coordSlots.size();
                it.rowB = (int)rowSlot_.size();
                for (int b = 0; b < nB; b++) {
                    int s = slot(c.atoms[1 + nA + b]);
                    coordSlots.push_back(s);
                    addRow(s, v);
                }
                it.r0 = (float)c.params[0];
                it.rn = (float)c.params[1];
                it.rm = (float)c.params[2];
                it.intPow = (it.rn == std::floor(it.rn) && it.rm == std::floor(it.rm) && it.rn >= 0 && it.rm >= 0 &&
                             it.rn <= 64 && it.rm <= 64);
                const double rc = c.params.size() > 3 ? c.params[3] : 0.0;
                it.rc2 = rc > 0.0 ? rc * rc : 0.0;
                {
                    // CommonGluedKernels.cpp:979-1007 (Taylor branch constants)
                    const double r0 = it.r0, rn = it.rn, rm = it.rm;
                    it.ir0 = 1.0 / r0;
                    it.preRes = rn / rm;
                    it.preDfunc = 0.5 * rn * (rn - rm) / rm;
                    it.preSecDev = rn * (rm * rm - 3.0 * rm * (rn - 1.0) + rn * (2.0 * rn - 3.0)) / (6.0 * rm);
                    it.ir0sq = it.ir0 * it.ir0;
                    it.cellBlock = tune_.coordBlock;
                    it.rat612 = rn == 6.0 && rm == 12.0;
                }
                coord_.push_back(it);
                break;
            }
            case CV_PATH: {
                PathItem it{};
                const int M = (int)c.atoms.size(), N = (int)c.params[1];
                it.M = M;
                it.N = N;
                it.lambda = c.params[0];
                it.aux = (int)path_.size();
                it.valueS = v;
                it.begin = (int)pathSlots.size();
                it.row0 = (int)rowSlot_.size();
                for (int j = 0; j < M; j++) {
                    int s = slot(c.atoms[j]);
                    pathSlots.push_back(s);
                    addRow(s, v);
                    addRow(s, v + 1);
                }
                it.frameOff = (int)pathFrames.size();
                it.msdOff = msdStride_;
                it.rowBlk0 = pathRowBlocks_;
                msdStride_ += N;
                pathRowBlocks_ += (M + PATH_TILE - 1) / PATH_TILE;
                {
                    it.centreOff = (int)pathCentreHost_.size();
                    pathCentreHost_.resize(pathCentreHost_.size() + (size_t)3 * M);
                    pathR2Host_.resize((size_t)msdStride_);
                }
                pathMaxN_ = std::max(pathMaxN_, N);
                for (int k = 0; k < N * M * 3; k++) pathFrames.push_back((float)c.params[2 + k]);
                path_.push_back(it);
                centrePath((int)path_.size() - 1, pathFrames.data() + it.frameOff);
                break;
            }
        }
    }
    // Slots relabelled in atom order.  The force scatters walk the slots in
    // order and read each slot's rows together, so atoms next to each other
    // in the force buffer (in the usual identity GPU order) then also have
    // their rows next to each other, and row reads share memory sectors.  A
    // relabelling only: every CV keeps its atom lists and rows, and each
    // slot's rows keep their order (the same sums, bit for bit).
    {
        const int S = (int)slotAtom_.size();
        std::vector<int> order(S), newOf(S), atoms(S);
        for (int k = 0; k < S; k++) order[k] = k;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return slotAtom_[a] < slotAtom_[b]; });
        for (int k = 0; k < S; k++) {
            newOf[order[k]] = k;
            atoms[k] = slotAtom_[order[k]];
        }
        slotAtom_ = atoms;
        for (auto& it : small_) {
            const int na = it.type == CV_DISTANCE ? 2 : it.type == CV_ANGLE ? 3 : it.type == CV_DIHEDRAL ? 4
                           : it.type == CV_POSITION ? 1 : 0;
            for (int a = 0; a < na; a++) it.slot[a] = newOf[it.slot[a]];
        }
        for (auto* v : {&groupSlots, &coordSlots, &pathSlots, &rowSlot_})
            for (int& x : *v) x = newOf[x];
    }
    S_ = (int)slotAtom_.size();
    R_ = (int)rowSlot_.size();
    // rows grouped by slot: the rows of slot s occupy [segStart_[s], segStart_[s + 1])
    segStart_.assign(S_ + 1, 0);
    for (int r = 0; r < R_; r++) segStart_[rowSlot_[r] + 1]++;
    for (int s = 0; s < S_; s++) segStart_[s + 1] += segStart_[s];

    // ---- walkers, atom order, box ----
    std::vector<WalkerInfo> wi(W_);
    for (int w = 0; w < W_; w++) {
        wi[w].posOff = posOff_[w];
        wi[w].forceOff = forceOff_[w];
        wi[w].padded = padded_[w];
        wi[w].natoms = natoms_[w];
    }
    dWalkers_.upload(wi);
    userToGpu_.assign(W_, std::vector<int>());
    if (posOff_[W_ - 1] + padded_[W_ - 1] > (long long)INT32_MAX) fail("initialize: more than 2^31 atoms in total");
    dGidx_.alloc(sizeof(int) * (size_t)W_ * std::max(S_, 1));
    groupSlots_ = groupSlots;
    pathSlots_ = pathSlots;
    coordSlots_ = coordSlots;
    dCoordPos_.alloc(sizeof(int) * (size_t)W_ * std::max((int)coordSlots_.size(), 1));
    dGroupPos_.alloc(sizeof(int) * (size_t)W_ * std::max((int)groupSlots_.size(), 1));
    dPathPos_.alloc(sizeof(int) * (size_t)W_ * std::max((int)pathSlots_.size(), 1));
    if (energyValue_ >= 0) {
        // GPU atom -> slot tables for the combined dense + slot scatter
        maskStride_ = (maxPadded_ + 31) / 32;
        dSlotMask_.alloc(sizeof(uint2) * (size_t)W_ * maskStride_);
        dSlotList_.alloc(sizeof(int2) * (size_t)W_ * std::max(S_, 1));
    } else {
        dSlotMask_.release();
        dSlotList_.release();
    }
    for (int w = 0; w < W_; w++) uploadGidx(w);
    boxHost_.assign((size_t)W_ * 12, 0.0);
    for (int w = 0; w < W_; w++) {
        double* b = &boxHost_[(size_t)w * 12];
        b[0] = b[4] = b[8] = 1.0;
        b[9] = b[10] = b[11] = 1.0;
    }
    dBox_.upload(boxHost_);

    // ---- per-walker step buffers ----
    dCv_.alloc(sizeof(double) * (size_t)W_ * std::max(V_, 1));
    dJ_.alloc(sizeof(float4) * (size_t)W_ * std::max(R_, 1));
    dUd_.alloc(sizeof(double) * (size_t)W_ * std::max(V_, 1));
    dSmall_.upload(small_);
    {
        // atom / pair lists longer than one 256-thread workgroup holds in
        // registers go to the 1,024-thread group kernel; contact maps longer
        // than one chunk are evaluated in chunks (partial sums + last chunk)
        std::vector<GroupCVItem> small, big;
        cm_.clear();
        cmBlocks_ = 0;
        for (GroupCVItem it : group_) {
            const bool pairs = it.type == CV_DRMSD || it.type == CV_CONTACTMAP;
            const int n = pairs ? (it.end - it.begin) / 2 : it.end - it.begin;
            if (it.type == CV_CONTACTMAP && n > CM_CHUNK) {
                it.chunk0 = cmBlocks_;
                it.chunkItem = (int)cm_.size();
                cmBlocks_ += (n + CM_CHUNK - 1) / CM_CHUNK;
                cm_.push_back(it);
                continue;
            }
            const bool fit = it.type == CV_RMSD && it.split;  // Kabsch stays on the 256-thread kernels
            (n > 1024 && !fit ? big : small).push_back(it);
        }
        group_ = small;
        groupBig_ = big;
        // their workgroup size: 512 threads when a path-row launch can carry
        // them (a plan with a path CV and no coordination, whose rank launch
        // carries 1,024-thread items), else 1,024
        groupBigNT_ = !path_.empty() && coord_.empty() ? 512 : 1024;
    }
    dGroup_.upload(group_);
    dGroupBig_.upload(groupBig_);
    dCm_.upload(cm_);
    dCmPart_.alloc(sizeof(double) * (size_t)W_ * std::max(cmBlocks_, 1));
    dCmCount_.alloc(sizeof(unsigned int) * (size_t)W_ * std::max((int)cm_.size(), 1));
    dGroupMass_.upload(groupMass);
    dGroupRef_.upload(groupRef);
    dPathAux_.alloc(sizeof(double) * (size_t)W_ * std::max((int)path_.size(), 1) * 2);
    dPathMsd_.alloc(sizeof(double) * (size_t)W_ * std::max(msdStride_, 1));
    dPathFrames_.upload(pathFrames);
    dPathCentre_.upload(pathCentreHost_);
    dPathR2_.upload(pathR2Host_);
    planPathWork();
    // row phases: one choice for every launch of this engine (one walker or
    // all), 4 when the batch's row blocks exceed one residency round and the
    // paths have at most 256 frames, else 8 (the autotuner may change it)
    if (tune_.pathPhases == 0) tune_.pathPhases = (long long)pathRowBlocks_ * W_ > 512 && pathMaxN_ <= 256 ? 4 : 8;
    {
        // the partial buffer holds the most slices any tuned shape can use
        long long most = 0;
        for (const auto& it : path_) most += (long long)PATH_MAX_SLICES * (it.N + 1);
        dPathPart_.alloc(sizeof(double) * (size_t)W_ * std::max<long long>(most, 1));
    }
    energyU_.assign(W_, 0.0);
    energyF_.assign(W_, nullptr);
    dEnergyU_.alloc(sizeof(double) * W_);
    dEnergyFPtr_.alloc(sizeof(const float*) * W_);

    // ---- coordination buffers ----
    {
        int cellBase = 0, sortBase = 0, cellOf = 0;
        for (auto& it : coord_) {
            it.useCells = 0;
            it.nc[0] = it.nc[1] = it.nc[2] = 1;
            it.cellBase = cellBase;
            it.sortBase = sortBase;
            cellBase += 2;
            sortBase += it.nA + it.nB;
            cellOf += it.nA + it.nB;
        }
        cellOfStride_ = std::max(cellOf, 1);
        sortStride_ = std::max(sortBase, 1);
        cellStride_ = std::max(cellBase, 1);
        const size_t nci = (size_t)W_ * std::max((int)coord_.size(), 1);
        dCellOf_.alloc(sizeof(int) * (size_t)W_ * cellOfStride_);
        dSorted_.alloc(sizeof(float4) * (size_t)W_ * sortStride_);
        dUnsorted_.alloc(sizeof(float4) * (size_t)W_ * sortStride_);
        dRank_.alloc(sizeof(int) * (size_t)W_ * sortStride_);
        layoutCellTables();
        dCnAcc_.alloc(sizeof(long long) * nci);
        dCnAccLo_.alloc(sizeof(unsigned long long) * nci);
        dPairAcc_.alloc(sizeof(unsigned long long) * nci);
        dNpairs_.alloc(sizeof(long long) * nci);
        dCoordTicket_.alloc(sizeof(unsigned int) * nci);
        dCellCount_.alloc(sizeof(int) * (size_t)W_ * cellStride_);
        dCellStart_.alloc(sizeof(int) * (size_t)W_ * (cellStride_ + coord_.size() + 1));
        dCoord_.upload(coord_);
        dCoordSlots_.upload(coordSlots);
    }

    // ---- biases ----
    int gslot = 0;
    std::vector<std::vector<int>> comb(V_);
    std::vector<GridBiasDev> gridTable;
    int scalCount = 0, varCount = 0;
    for (size_t b = 0; b < biases_.size(); b++) {
        const BiasSpec& s = biases_[b];
        const int D = (int)s.cvs.size();
        const int T = s.shared ? 1 : W_;  // tables
        for (int d = 0; d < D; d++) comb[s.cvs[d]].push_back(gslot + d);
        if (s.type == BIAS_METAD) {
            auto* g = new GridHost();
            GridBiasDev& d = g->d;
            d = GridBiasDev{};
            d.D = D;
            for (int k = 0; k < D; k++) d.cv[k] = s.cvs[k];
            d.gslot = gslot;
            d.eslot = (int)b;
            d.h0 = s.params[0];
            for (int k = 0; k < D; k++) d.sigma[k] = s.params[1 + k];
            double gamma = s.params[1 + D], kT = s.params[2 + D];
            for (int k = 0; k < D; k++) {
                double mn = s.params[3 + D + k], mx = s.params[3 + 2 * D + k];
                int bins = s.iparams[1 + k];
                d.periodic[k] = s.iparams[1 + D + k];
                if (bins < 1 || !(mx > mn)) fail("METAD: invalid grid range or bin count");
                d.origin[k] = mn;
                d.invSpacing[k] = bins / (mx - mn);
                d.npts[k] = bins + (d.periodic[k] ? 0 : 1);
            }
            d.strides[0] = 1;
            for (int k = 1; k < D; k++) d.strides[k] = d.strides[k - 1] * d.npts[k - 1];
            d.total = d.strides[D - 1] * d.npts[D - 1];
            d.hif = gamma <= 1.0 ? 0.0 : 1.0 / ((gamma - 1.0) * kT);
            g->pace = s.iparams[0];
            g->spec = (int)b;
            g->numDeposited.assign(T, 0);
            g->grid.alloc(sizeof(double) * (size_t)T * d.total);
            d.grid = g->grid.as<double>();
            d.walkerStride = s.shared ? 0 : d.total;
            SmallBiasItem it{};
            it.type = BIAS_METAD;
            it.D = D;
            it.gslot = gslot;
            it.eslot = (int)b;
            it.bias = (int)gridTable.size();
            it.grid = d;
            gridTable.push_back(d);
            grids_.push_back(g);
            smallBias_.push_back(it);
            smallBiasHost_.push_back({(int)b, g->pace, BIAS_METAD});
        } else if (s.type == BIAS_METAD_HILLS) {
            auto* h = new HillHost();
            HillBiasDev& d = h->d;
            d = HillBiasDev{};
            d.D = D;
            for (int k = 0; k < D; k++) d.cv[k] = s.cvs[k];
            d.gslot = gslot;
            d.eslot = (int)b;
            d.h0 = s.params[0];
            double gamma = s.params[1 + D], kT = s.params[2 + D];
            for (int k = 0; k < D; k++) {
                d.sigma[k] = s.params[1 + k];
                d.period[k] = s.params[3 + D + k];
                if (!(d.sigma[k] > 0.0)) fail("hill METAD: sigma must be positive");
                d.i2s2[k] = 1.0 / (2.0 * d.sigma[k] * d.sigma[k]);
                d.is2[k] = 1.0 / (d.sigma[k] * d.sigma[k]);
                d.invPeriod[k] = d.period[k] > 0.0 ? 1.0 / d.period[k] : 0.0;
            }
            d.hif = gamma <= 1.0 ? 0.0 : 1.0 / ((gamma - 1.0) * kT);
            d.cap = std::max(1, s.iparams[1]);
            d.shared = s.shared;
            d.chunks = hill_chunks_for(d.cap);
            h->pace = s.iparams[0];
            h->spec = (int)b;
            h->centers.alloc(sizeof(float) * (size_t)T * D * d.cap);
            h->heights.alloc(sizeof(float) * (size_t)T * d.cap);
            h->count.alloc(sizeof(int) * T);
            h->part.alloc(sizeof(double) * (size_t)W_ * d.chunks * (D + 1));
            h->counter.alloc(sizeof(unsigned int) * W_);
            d.centers = h->centers.as<float>();
            d.heights = h->heights.as<float>();
            d.count = h->count.as<int>();
            d.part = h->part.as<double>();
            d.counter = h->counter.as<unsigned int>();
            hills_.push_back(h);
        } else if (s.type == BIAS_OPES) {
            auto* o = new OpesHost();
            OpesBiasDev& d = o->d;
            d = OpesBiasDev{};
            d.D = D;
            for (int k = 0; k < D; k++) d.cv[k] = s.cvs[k];
            d.gslot = gslot;
            d.eslot = (int)b;
            d.kT = s.params[0];
            d.gamma = s.params[1];
            for (int k = 0; k < D; k++) d.sigma0[k] = s.params[2 + k];
            d.sigmaMin = s.params[2 + D];
            d.variant = s.iparams.size() > 0 ? s.iparams[0] : 0;
            o->pace = s.iparams.size() > 1 ? s.iparams[1] : 500;
            d.cap = s.iparams.size() > 2 ? s.iparams[2] : 100000;
            const bool adaptive = d.sigma0[0] <= 0.0;
            d.adaptiveStride = adaptive ? (s.iparams.size() > 3 ? s.iparams[3] : 10 * o->pace) : 0;
            // variant constants, CommonBiasKernels.cpp:553-588
            if (d.variant == 1) {
                d.invGF = 1.0;
                d.cutoff2 = 2.0 * d.gamma;
                d.eps = 0.0;
                d.barrier = d.gamma * d.kT;
            } else if (d.variant == 2) {
                if (!std::isfinite(d.gamma) || d.gamma <= 1.0)
                    fail("OPES_METAD_EXPLORE (variant=2) requires finite gamma > 1");
                d.invGF = d.gamma - 1.0;
                d.cutoff2 = 2.0 * d.gamma;
                d.eps = std::exp(-d.gamma / (d.gamma - 1.0));
                d.barrier = d.gamma * d.kT;
                if (!adaptive)
                    for (int k = 0; k < D; k++) d.sigma0[k] *= std::sqrt(d.gamma);
            } else {
                d.invGF = (d.gamma - 1.0) / d.gamma;
                d.cutoff2 = 2.0 * d.gamma / d.invGF;
                d.eps = std::exp(-1.0 / (d.invGF * (1.0 - d.invGF)));
                d.barrier = d.kT / (1.0 - d.invGF);
            }
            d.shared = s.shared;
            d.chunks = opes_chunks_for(d.cap);
            o->spec = (int)b;
            o->depositMirror.assign(T, 0);
            o->nSamplesMirror.assign(T, 0);
            o->centers.alloc(sizeof(double) * (size_t)T * D * d.cap);
            o->sigmas.alloc(sizeof(double) * (size_t)T * D * d.cap);
            o->logw.alloc(sizeof(double) * (size_t)T * d.cap);
            o->ints.alloc(sizeof(int) * 4 * T);
            std::vector<double> dbl((size_t)T * (4 + 2 * D), 0.0);
            for (int t = 0; t < T; t++) {
                dbl[(size_t)t * (4 + 2 * D) + 1] = std::exp(-d.gamma);        // KDNorm sentinel
                dbl[(size_t)t * (4 + 2 * D) + 2] = std::exp(-d.gamma);        // sum_w sentinel
                dbl[(size_t)t * (4 + 2 * D) + 3] = std::exp(-2.0 * d.gamma);  // sum_w2 sentinel
            }
            o->dbl.upload(dbl);
            o->part.alloc(sizeof(double) * (size_t)W_ * d.chunks * (D + 1));
            o->counter.alloc(sizeof(unsigned int) * W_);
            d.centers = o->centers.as<double>();
            d.sigmas = o->sigmas.as<double>();
            d.logw = o->logw.as<double>();
            d.ints = o->ints.as<int>();
            d.dbl = o->dbl.as<double>();
            d.part = o->part.as<double>();
            d.counter = o->counter.as<unsigned int>();
            // two-launch deposit of private tables: the workgroups' partials,
            // the change record and two tickets per table
            o->depPart.alloc(sizeof(double) * (size_t)T * OPES_DEP_G * OPES_PART);
            o->depRec.alloc(sizeof(double) * (size_t)T * OPES_REC);
            o->depCount.alloc(sizeof(unsigned int) * 2 * T);
            d.depPart = o->depPart.as<double>();
            d.depRec = o->depRec.as<double>();
            d.depCount = o->depCount.as<unsigned int>();
            opes_.push_back(o);
        } else if (s.type == BIAS_OPES_EXPANDED || s.type == BIAS_OPES_MULTITHERMAL) {
            SmallBiasItem it{};
            it.type = s.type;
            it.gslot = gslot;
            it.eslot = (int)b;
            it.cv[0] = s.cvs[0];
            it.stateOff = (int)smallConst_.size();
            int pace = s.iparams.size() > 0 ? s.iparams[0] : 500;
            if (s.type == BIAS_OPES_EXPANDED) {
                it.D = D;
                it.nState = D;
                it.p[0] = 1.0 / s.params[0];
                double ws = 0.0;
                for (int l = 0; l < D; l++) ws += s.params[1 + l];
                if (ws <= 0.0) ws = D;
                for (int l = 0; l < D; l++) smallConst_.push_back(std::log(s.params[1 + l] / ws));
                // the state's CV value index (any order, CommonBiasKernels.cpp:740-742)
                for (int l = 0; l < D; l++) smallConst_.push_back((double)s.cvs[l]);
            } else {
                const int N = (int)s.params.size() - 1;
                it.D = 1;
                it.nState = N;
                it.p[0] = s.params[0];
                const double beta0 = 1.0 / s.params[0];
                for (int l = 0; l < N; l++) smallConst_.push_back(s.params[1 + l] - beta0);
                varCount = std::max(varCount, it.stateOff + N);
            }
            scalCount++;
            smallBias_.push_back(it);
            smallBiasHost_.push_back({(int)b, pace, s.type});
        } else if (s.type == BIAS_PBMETAD) {
            // CommonBiasKernels.cpp:298-400: one 1-D grid per CV,
            // spacing (max - min) / bins, points bins (+1 unless periodic)
            SmallBiasItem it{};
            it.type = s.type;
            it.gslot = gslot;
            it.eslot = (int)b;
            it.D = D;
            for (int d = 0; d < D; d++) it.cv[d] = s.cvs[d];
            const double gamma = s.params[1], kT = s.params[2];
            it.p[0] = s.params[0];
            it.p[1] = kT;
            it.p[2] = gamma <= 1.0 ? 0.0 : 1.0 / ((gamma - 1.0) * kT);
            it.p[3] = s.shared ? 1.0 : 0.0;
            it.stateOff = (int)smallConst_.size();
            PbHost ph;
            ph.item = (int)smallBias_.size();
            ph.N = D;
            ph.T = s.shared ? 1 : W_;
            ph.pace = s.iparams[0];
            long long sub = 0;
            for (int d = 0; d < D; d++) {
                const int bins = s.iparams[1 + 2 * d], per = s.iparams[2 + 2 * d] != 0;
                const double sigma = s.params[3 + 3 * d], lo = s.params[4 + 3 * d], hi = s.params[5 + 3 * d];
                const int np = bins + (per ? 0 : 1);
                for (double x : {lo, bins / (hi - lo), (double)np, (double)per, sigma, (double)sub})
                    smallConst_.push_back(x);
                ph.npts.push_back(np);
                sub += np;
            }
            ph.stride = sub;
            ph.base = pbGridTotal_;
            pbGridTotal_ += ph.T * sub;
            if (pbGridTotal_ > (long long)INT32_MAX) fail("PBMetaD grids too large");
            it.bias = (int)ph.base;
            it.nState = (int)sub;
            ph.numDeposited.assign(ph.T, 0);
            pbs_.push_back(ph);
            smallBias_.push_back(it);
            smallBiasHost_.push_back({(int)b, ph.pace, s.type});
        } else if (s.type == BIAS_HARMONIC || s.type == BIAS_LINEAR || s.type == BIAS_UPPER_WALL ||
                   s.type == BIAS_LOWER_WALL) {
            // static restraints: parameters rounded to FP32 as the reference
            // stores them (CommonBiasKernels.cpp:118-126, 676-678, 700-707)
            SmallBiasItem it{};
            it.type = s.type;
            it.gslot = gslot;
            it.eslot = (int)b;
            it.D = D;
            for (int d = 0; d < D; d++) it.cv[d] = s.cvs[d];
            const int per = s.type == BIAS_HARMONIC ? 2 : s.type == BIAS_LINEAR ? 1 : 4;
            it.stateOff = (int)smallConst_.size();
            for (int k = 0; k < per * D; k++) smallConst_.push_back((double)(float)s.params[k]);
            it.bias = s.type == BIAS_HARMONIC ? (s.iparams.empty() ? 0 : s.iparams[0] != 0)
                                              : (s.type == BIAS_LOWER_WALL ? 1 : 0);
            smallBias_.push_back(it);
            smallBiasHost_.push_back({(int)b, 0, s.type});
        }
        gslot += D;
    }
    G_ = std::max(gslot, 1);
    B_ = std::max((int)biases_.size(), 1);
    dBgrad_.alloc(sizeof(double) * (size_t)W_ * G_);
    dBener_.alloc(sizeof(double) * (size_t)W_ * B_);
    dGridTable_.upload(gridTable);
    dSmallBias_.upload(smallBias_);
    dSmallConst_.upload(smallConst_);
    dPbGrid_.alloc(sizeof(double) * (size_t)std::max(pbGridTotal_, 1LL));
    smallVarPerWalker_ = std::max((int)smallConst_.size(), 1);
    dSmallVar_.alloc(sizeof(double) * (size_t)W_ * smallVarPerWalker_);
    {
        // per walker per small bias: {logZ, numUpdates, rct, counter}; multithermal counter starts at 1
        std::vector<double> scal((size_t)W_ * std::max((int)smallBias_.size(), 1) * 4, 0.0);
        for (int w = 0; w < W_; w++)
            for (size_t i = 0; i < smallBias_.size(); i++)
                if (smallBias_[i].type == BIAS_OPES_MULTITHERMAL) scal[((size_t)w * smallBias_.size() + i) * 4 + 3] = 1.0;
        dSmallScal_.upload(scal);
    }
    combStart_.assign(V_ + 1, 0);
    for (int v = 0; v < V_; v++) {
        combStart_[v + 1] = combStart_[v] + (int)comb[v].size();
        for (int g : comb[v]) combSlot_.push_back(g);
    }
    dCombStart_.upload(combStart_);
    dCombSlot_.upload(combSlot_);
    if (!testGrads_.empty()) {
        if ((int)testGrads_.size() != V_) fail("setTestBiasGradients: length must equal the number of CV values");
        dTestGrad_.upload(testGrads_);
    }

    // ---- scatter segments: rows grouped by atom slot ----
    // (row r is stored at rowPos[r]; the rows of slot s occupy [segStart[s], segStart[s+1]))
    const std::vector<int>& segStart = segStart_;
    std::vector<int> rowPos(R_), entryCv(R_);
    {
        std::vector<int> cur(segStart.begin(), segStart.end() - 1);
        for (int r = 0; r < R_; r++) {
            const int k = cur[rowSlot_[r]]++;
            rowPos[r] = k;
            entryCv[k] = rowCv_[r];
        }
    }
    dSegStart_.upload(segStart);
    dSegRow_.upload(rowPos);
    dRowCv_.upload(entryCv);

    // ---- the batched bias + scatter stage ----
    {
        std::vector<HillBiasDev> ht;
        std::vector<OpesBiasDev> ot;
        for (auto* h : hills_) ht.push_back(h->d);
        for (auto* o : opes_) ot.push_back(o->d);
        if (ht.size() + ot.size() > (size_t)MAX_CHUNKED) fail("too many hill / OPES biases (max 16)");
        dHillTable_.upload(ht);
        dOpesTable_.upload(ot);
        fplan_ = FusedBiasPlan{};
        fplan_.hills = dHillTable_.as<HillBiasDev>();
        if (!ht.empty()) fplan_.hill0 = ht[0];
        // one shared hill list and nothing else chunked: the chunk
        // workgroups take CHUNK_WB walkers each
        fplan_.walkersPerChunk = (ht.size() == 1 && ot.empty() && ht[0].shared && W_ > 1) ? CHUNK_WB : 1;
        fplan_.opes = dOpesTable_.as<OpesBiasDev>();
        fplan_.nHill = (int)ht.size();
        fplan_.nOpes = (int)ot.size();
        fplan_.nChunked = fplan_.nHill + fplan_.nOpes;
        int c = 0;
        for (int k = 0; k < fplan_.nChunked; k++) {
            fplan_.chunkStart[k] = c;
            c += k < fplan_.nHill ? ht[k].chunks : ot[k - fplan_.nHill].chunks;
        }
        fplan_.chunkStart[fplan_.nChunked] = c;
        fplan_.totalChunks = c;
        dFusedPart_.alloc(sizeof(double) * (size_t)W_ * std::max(c, 1) * (MAX_BIAS_D + 1));
        dFusedCounter_.alloc(sizeof(unsigned int) * W_);
        fplan_.part = dFusedPart_.as<double>();
        fplan_.counter = dFusedCounter_.as<unsigned int>();
        fplan_.small = dSmallBias_.as<SmallBiasItem>();
        fplan_.nSmall = (int)smallBias_.size();
        fplan_.extra = 0;
        for (auto& it : smallBias_)
            fplan_.extra |= it.type == BIAS_HARMONIC || it.type == BIAS_LINEAR || it.type == BIAS_UPPER_WALL ||
                            it.type == BIAS_LOWER_WALL || it.type == BIAS_PBMETAD;
        // one workgroup scatters a walker when its atom list is short
        fplan_.scatter = (S_ <= 4096 && V_ <= 1024) ? 1 : 0;
        // all chunked biases of one type and dimension: a specialised chunk
        // kernel (fewer registers, more resident workgroups)
        fplan_.kind = 0;
        if (!ht.empty() && ot.empty()) {
            fplan_.kind = 1;
            fplan_.kindD = ht[0].D;
            for (auto& h : ht) fplan_.kind = h.D == fplan_.kindD ? fplan_.kind : 0;
        } else if (ht.empty() && !ot.empty()) {
            fplan_.kind = 2;
            fplan_.kindD = ot[0].D;
            for (auto& o : ot) fplan_.kind = o.D == fplan_.kindD ? fplan_.kind : 0;
        }
        if (fplan_.kind && (fplan_.kindD < 1 || fplan_.kindD > 3)) fplan_.kind = 0;
    }
}

// MSD work of every path: tiles of PATH_TW walkers x PATH_TF frames x K
// slices, about tune_.pathBlocks blocks per path over the batch, at most
// PATH_MAX_SLICES slices; the per-kind launch takes one unit per block and
// walker tile, the fused CV launch spreads (walker tile, unit) items over
// msdBlkF blocks of every walker row.
void Engine::planPathWork()
{
    pathMsdUnits_ = pathMsdBlocks_ = partStride_ = 0;
    const int wTiles = (W_ + PATH_TW - 1) / PATH_TW;
    for (auto& it : path_) {
        const int nChunks = (3 * it.M + PATH_KC - 1) / PATH_KC;
        it.frameTiles = (it.N + PATH_TF - 1) / PATH_TF;
        int sl = (tune_.pathBlocks + wTiles * it.frameTiles - 1) / (wTiles * it.frameTiles);
        sl = std::max(1, std::min(sl, std::min(nChunks, PATH_MAX_SLICES)));
        it.chunksPerSlice = (nChunks + sl - 1) / sl;
        it.slices = (nChunks + it.chunksPerSlice - 1) / it.chunksPerSlice;
        it.units = it.frameTiles * it.slices;
        it.msdBlk0 = pathMsdUnits_;
        pathMsdUnits_ += it.units;
        it.msdBlkF0 = pathMsdBlocks_;
        it.msdBlkF = (int)(((long long)wTiles * it.units + W_ - 1) / W_);
        pathMsdBlocks_ += it.msdBlkF;
        it.partOff = partStride_;
        partStride_ += it.slices * it.N + it.slices;
    }
    dPath_.upload(path_);
}

// ---------------------------------------------------------------------------
// Launch-shape autotuning.  Candidates: the path MSD block target
// {128, 256, 512, 1024}, the path row phases {4, 8} and the coordination cell
// workgroup {256, 512, 1024}.  At an engine's first evaluation, when it is a
// batched one, each candidate's CV stage (every walker) is timed with HIP
// events on the caller's stream (3 launches to warm, the mean of 8), one
// group at a time, the others at their current best; the choice is cached
// per plan shape for the process, so later engines of the same shape take
// it without timing.  An engine whose first evaluation is a one-walker
// launch keeps the defaults (or a cached choice), so that all of its
// evaluations, plain or batched, run the same shapes and give the same bits.
// GLUED_AUTOTUNE=0 in the environment keeps the defaults.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tuneMu;
std::map<std::string, Engine::Tuning> g_tuneCache;
}  // namespace

std::string Engine::shapeKey() const
{
    std::string k = "W" + std::to_string(W_);
    for (int n : natoms_) k += "," + std::to_string(n);
    for (const auto& it : path_) k += ";path" + std::to_string(it.N) + "x" + std::to_string(it.M);
    for (const auto& it : coord_)
        k += ";coord" + std::to_string(it.nA) + "x" + std::to_string(it.nB) + ":" + std::to_string(it.useCells) + ":" +
             std::to_string(it.nc[0]) + "," + std::to_string(it.nc[1]) + "," + std::to_string(it.nc[2]);
    return k;
}

void Engine::applyTuning()
{
    if (!path_.empty()) planPathWork();
    if (!coord_.empty()) {
        for (auto& it : coord_) it.cellBlock = tune_.coordBlock;
        dCoord_.upload(coord_);
    }
    dropGraphs();
}

void Engine::adoptCachedTuning()
{
    std::lock_guard<std::mutex> lock(g_tuneMu);
    auto f = g_tuneCache.find(shapeKey());
    if (f == g_tuneCache.end()) return;
    tune_ = f->second;
    tuneHow_ = "cached";
    applyTuning();
}

void Engine::autotune(hipStream_t s)
{
    if (tuned_) return;
    tuned_ = true;
    if (evaluated_ || capturing_) return;
    {
        // under the caller's own graph capture nothing can be timed: keep the defaults
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    }
    const char* env = std::getenv("GLUED_AUTOTUNE");
    if (env && env[0] == '0') return;
    bool cells = false;
    for (const auto& it : coord_) cells |= it.useCells != 0;
    if (path_.empty() && !cells) return;
    adoptCachedTuning();
    if (tuneHow_ == "cached") return;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    auto timeCv = [&]() {
        applyTuning();
        for (int r = 0; r < 3; r++) launchStep(s, 0, W_, true, 1, true);
        hip_check(hipEventRecord(e0, s), "event");
        for (int r = 0; r < 8; r++) launchStep(s, 0, W_, true, 1, true);
        hip_check(hipEventRecord(e1, s), "event");
        hip_check(hipEventSynchronize(e1), "event");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
        return 1000.0 * ms / 8.0;  // us per CV stage
    };
    std::string log;
    auto search = [&](int Tuning::*field, std::initializer_list<int> values, const char* name) {
        int best = tune_.*field;
        double bestT = 1e300;
        log += std::string(name) + ":";
        for (int val : values) {
            tune_.*field = val;
            const double t = timeCv();
            char b[48];
            std::snprintf(b, sizeof(b), " %d=%.1fus", val, t);
            log += b;
            if (t < bestT) {
                bestT = t;
                best = val;
            }
        }
        tune_.*field = best;
        log += "; ";
    };
    if (cells) search(&Tuning::coordBlock, {256, 512, 1024}, "coordBlock");
    if (!path_.empty()) {
        search(&Tuning::pathBlocks, {128, 256, 512, 1024}, "pathBlocks");
        search(&Tuning::pathPhases, {4, 8}, "pathPhases");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    applyTuning();
    tuneHow_ = "measured";
    tuneLog_ = log;
    std::lock_guard<std::mutex> lock(g_tuneMu);
    g_tuneCache[shapeKey()] = tune_;
}

std::string Engine::tuningReport() const
{
    return "pathBlocks=" + std::to_string(tune_.pathBlocks) + " pathPhases=" + std::to_string(tune_.pathPhases) +
           " coordBlock=" + std::to_string(tune_.coordBlock) + " (" + tuneHow_ + ")" +
           (tuneLog_.empty() ? "" : " " + tuneLog_);
}

void Engine::uploadGidx(int w)
{
    std::vector<int> g(std::max(S_, 1), 0);
    for (int s = 0; s < S_; s++) {
        int u = slotAtom_[s];
        g[s] = userToGpu_[w].empty() ? u : userToGpu_[w][u];
    }
    hip_check(hipMemcpy(dGidx_.as<int>() + (size_t)w * std::max(S_, 1), g.data(), sizeof(int) * g.size(),
                        hipMemcpyHostToDevice),
              "upload atom indices");
    // resolved gather tables: posq index of every group / path entry of this walker
    auto resolve = [&](const std::vector<int>& slots, DBuf& dst) {
        if (slots.empty()) return;
        std::vector<int> p(slots.size());
        for (size_t e = 0; e < slots.size(); e++) p[e] = (int)posOff_[w] + g[slots[e]];
        hip_check(hipMemcpy(dst.as<int>() + (size_t)w * slots.size(), p.data(), sizeof(int) * p.size(),
                            hipMemcpyHostToDevice),
                  "upload gather table");
    };
    resolve(groupSlots_, dGroupPos_);
    resolve(pathSlots_, dPathPos_);
    resolve(coordSlots_, dCoordPos_);
    if (dSlotMask_.bytes) {
        // per 32-atom word: the bits of the atoms with Jacobian rows and the
        // count of such atoms in the earlier words; the row ranges of those
        // atoms in GPU order
        std::vector<int> so((size_t)maskStride_ * 32, -1);
        for (int s = 0; s < S_; s++) so[g[s]] = s;
        std::vector<uint2> mk((size_t)maskStride_);
        std::vector<int2> sl;
        sl.reserve(std::max(S_, 1));
        for (int wd = 0; wd < maskStride_; wd++) {
            unsigned bits = 0;
            const unsigned before = (unsigned)sl.size();
            for (int b = 0; b < 32; b++) {
                const int s = so[(size_t)wd * 32 + b];
                if (s >= 0) {
                    bits |= 1u << b;
                    sl.push_back(make_int2(segStart_[s], segStart_[s + 1]));
                }
            }
            mk[wd] = make_uint2(bits, before);
        }
        sl.resize(std::max(S_, 1), make_int2(0, 0));
        hip_check(hipMemcpy(dSlotMask_.as<uint2>() + (size_t)w * maskStride_, mk.data(), sizeof(uint2) * mk.size(),
                            hipMemcpyHostToDevice),
                  "upload slot mask");
        hip_check(hipMemcpy(dSlotList_.as<int2>() + (size_t)w * std::max(S_, 1), sl.data(), sizeof(int2) * sl.size(),
                            hipMemcpyHostToDevice),
                  "upload slot list");
    }
}

// rebuildGpuAtomIndices, CommonCVKernels.cpp:113-317: user -> GPU order.
void Engine::setAtomOrder(int w, const int* gpuToUser, int n)
{
    if (!initialized_) fail("set_atom_order: initialize first");
    if (w < 0 || w >= W_) fail("set_atom_order: walker out of range");
    if (n < natoms_[w]) fail("set_atom_order: array shorter than the atom count");
    std::vector<int> u2g(natoms_[w], -1);
    for (int g = 0; g < n; g++) {
        int u = gpuToUser[g];
        if (u >= 0 && u < natoms_[w]) u2g[u] = g;
    }
    for (int s = 0; s < S_; s++)
        if (u2g[slotAtom_[s]] < 0) fail("set_atom_order: atom " + std::to_string(slotAtom_[s]) + " missing");
    userToGpu_[w] = u2g;
    uploadGidx(w);
}

void Engine::setBox(int w, const double* b9)
{
    if (!initialized_) fail("set_box: initialize first");
    if (w < 0 || w >= W_) fail("set_box: walker out of range");
    double b[12];
    for (int k = 0; k < 9; k++) b[k] = b9[k];
    if (!(b[0] > 0 && b[4] > 0 && b[8] > 0)) fail("set_box: box vectors must have positive diagonal");
    b[9] = 1.0 / b[0];
    b[10] = 1.0 / b[4];
    b[11] = 1.0 / b[8];
    if (std::memcmp(b, &boxHost_[(size_t)w * 12], sizeof(b)) == 0) return;
    std::memcpy(&boxHost_[(size_t)w * 12], b, sizeof(b));
    hip_check(hipMemcpy(dBox_.as<double>() + (size_t)w * 12, b, sizeof(b), hipMemcpyHostToDevice), "upload box");
    // The cell grid of a coordination item is shared by every walker: it is
    // used only when every walker's box is orthorhombic, and its per-axis
    // cell count is the smallest over the walkers, so that every walker's
    // cell edge is at least the cutoff (the pair passes visit the +-1
    // neighbour cells only).  Recomputed over all boxes on every change.
    bool changed = false;
    for (size_t i = 0; i < coord_.size(); i++) {
        CoordItem& it = coord_[i];
        int nc[3] = {64, 64, 64};
        int use = pbc && it.rc2 > 0.0;
        const double rc = std::sqrt(it.rc2) * (1.0 + 1e-6);
        for (int v = 0; v < W_ && use; v++) {
            const double* bv = &boxHost_[(size_t)v * 12];
            const bool ortho = bv[1] == 0 && bv[2] == 0 && bv[3] == 0 && bv[5] == 0 && bv[6] == 0 && bv[7] == 0;
            if (!ortho) use = 0;
            for (int k = 0; k < 3 && use; k++) {
                const int n = (int)std::floor(bv[4 * k] / rc);
                if (n < 3) use = 0;
                nc[k] = std::min(nc[k], n);
            }
        }
        if (!use) nc[0] = nc[1] = nc[2] = 1;
        if (use != it.useCells || nc[0] != it.nc[0] || nc[1] != it.nc[1] || nc[2] != it.nc[2]) changed = true;
        it.useCells = use;
        for (int k = 0; k < 3; k++) it.nc[k] = nc[k];
    }
    if (changed) {
        int cellBase = 0;
        for (auto& it : coord_) {
            it.cellBase = cellBase;
            cellBase += 2 * it.nc[0] * it.nc[1] * it.nc[2];
        }
        if (cellBase > cellStride_) {
            cellStride_ = cellBase;
            dCellCount_.alloc(sizeof(int) * (size_t)W_ * cellStride_);
            dCellStart_.alloc(sizeof(int) * (size_t)W_ * (cellStride_ + coord_.size() + 1));
        }
        layoutCellTables();
        dCoord_.upload(coord_);
        dropGraphs();  // the coordination launches are shaped by the cell grid
    }
}

// The counting sort's (cell, workgroup) table: 2 * ncell * ceil(n / 1024)
// entries per item and walker.
void Engine::layoutCellTables()
{
    int blk = 0;
    for (auto& it : coord_) {
        it.blkBase = blk;
        blk += 2 * it.nc[0] * it.nc[1] * it.nc[2] * ((it.nA + it.nB + 1023) / 1024);
    }
    if (std::max(blk, 1) > blkStride_) {
        blkStride_ = std::max(blk, 1);
        dBlockCnt_.alloc(sizeof(int) * (size_t)W_ * blkStride_);
    }
}

// CV_ENERGY input: U and the external force F (device, field-major FP32 with
// the walker's force-block layout).  The energy CV's Jacobian is -F.
void Engine::setEnergyInput(int w, double U, const float* Fext, hipStream_t s)
{
    if (!initialized_) fail("set_energy_input: initialize first");
    if (w < 0 || w >= W_) fail("set_energy_input: walker out of range");
    if (energyValue_ < 0) fail("set_energy_input: the force has no CV_ENERGY");
    if (Fext && (reinterpret_cast<uintptr_t>(Fext) & 15)) fail("set_energy_input: force array must be 16-byte aligned");
    energyU_[w] = U;
    energyF_[w] = Fext;
    hip_check(hipMemcpyAsync(dEnergyU_.as<double>() + w, &energyU_[w], sizeof(double), hipMemcpyHostToDevice, s),
              "energy input");
    hip_check(hipMemcpyAsync(dEnergyFPtr_.as<const float*>() + w, &energyF_[w], sizeof(const float*),
                             hipMemcpyHostToDevice, s),
              "energy input");
    hip_check(hipStreamSynchronize(s), "energy input sync");
}

StepView Engine::view(const float4* posq, long long* force) const
{
    StepView v;
    v.posq = posq;
    v.force = force;
    v.walkers = dWalkers_.as<WalkerInfo>();
    v.gidx = dGidx_.as<int>();
    v.box = dBox_.as<double>();
    v.cvVal = dCv_.as<double>();
    v.rowPos = dSegRow_.as<int>();
    v.jrow = dJ_.as<float4>();
    v.bgrad = dBgrad_.as<double>();
    v.benergy = dBener_.as<double>();
    v.dUdCV = dUd_.as<double>();
    v.energyU = dEnergyU_.as<double>();
    v.pathAux = dPathAux_.as<double>();
    v.nPath = std::max((int)path_.size(), 1);
    v.pathMsd = dPathMsd_.as<double>();
    v.msdStride = std::max(msdStride_, 1);
    v.pathPart = dPathPart_.as<double>();
    v.partStride = std::max(partStride_, 1);
    v.pathCentre = dPathCentre_.as<float>();
    v.pathR2 = dPathR2_.as<double>();
    v.gStride = (int)groupSlots_.size();
    v.pStride = (int)pathSlots_.size();
    v.S = std::max(S_, 1);
    v.V = std::max(V_, 1);
    v.R = std::max(R_, 1);
    v.G = G_;
    v.B = B_;
    v.pbc = pbc ? 1 : 0;
    // the kernels index with the padded strides used for allocation
    v.S = std::max(S_, 1);
    return v;
}

// ---------------------------------------------------------------------------
// One step for walkers [w0, w0+nw): CVs by kind, then biases, then the
// scatter (the ordering contract of CommonGluedKernels.cpp:4646-5048).
// ---------------------------------------------------------------------------
// Plans small enough for the whole batched step to run in one workgroup per
// walker: few-atom CVs only, no dense energy CV, the finalize's LDS limits,
// and at most 16 chunks of streamed bias state per walker.
bool Engine::tinyPlan() const
{
    return group_.empty() && groupBig_.empty() && cm_.empty() && path_.empty() && coord_.empty() && testGrads_.empty() &&
           energyValue_ < 0 && fplan_.scatter && V_ <= 1024 && G_ <= 512 && combSlot_.size() <= 1024 &&
           fplan_.totalChunks <= 16;
}

void Engine::launchStep(hipStream_t s, int w0, int nw, bool forces, int stageMask, bool fused)
{
    StepView v = view(lastPos_, lastForce_);
    v.S = S_;
    if (fused && stageMask == 7 && w0 == 0 && nw == W_ && tinyPlan()) {
        // the whole step in one launch (CVs, chunks, finalize and scatter per workgroup)
        FusedBiasPlan fp = fplan_;
        fp.st.grids = dGridTable_.as<GridBiasDev>();
        fp.st.stateConst = dSmallConst_.as<double>();
        fp.st.stateVar = dSmallVar_.as<double>();
        fp.st.scal = dSmallScal_.as<double>();
        fp.st.stateVarStride = smallVarPerWalker_;
        fp.st.scalStride = 4 * (long long)std::max((int)smallBias_.size(), 1);
        fp.st.pbGrid = dPbGrid_.as<double>();
        fp.sp = scatterPlan();
        fp.scatter = forces ? 1 : 0;
        launch_step_tiny(v, dSmall_.as<SmallCVItem>(), (int)small_.size(), fp, W_, s);
        // the every-step Welford update reads only the CV values
        for (auto* o : opes_)
            if (o->d.adaptiveStride > 0) {
                launch_opes_welford(v, o->d, 0, W_, s);
                if (!capturing_) bumpWelfordMirrors(o, 0, W_);
            }
        hip_check(hipGetLastError(), "kernel launch");
        return;
    }
    if (stageMask & 1) {
    CVFusedPlan cp;
    cp.small = dSmall_.as<SmallCVItem>();
    cp.nSmall = (int)small_.size();
    cp.group = dGroup_.as<GroupCVItem>();
    cp.nGroup = (int)group_.size();
    cp.gslots = dGroupPos_.as<int>();
    cp.gmass = dGroupMass_.as<float>();
    cp.gref = dGroupRef_.as<float>();
    cp.path = dPath_.as<PathItem>();
    cp.nPath = (int)path_.size();
    cp.nPathMsdBlocks = pathMsdBlocks_;
    cp.pslots = dPathPos_.as<int>();
    cp.frames = dPathFrames_.as<float>();
    cp.cm = dCm_.as<GroupCVItem>();
    cp.nCm = (int)cm_.size();
    cp.nCmBlocks = cmBlocks_;
    cp.cmPart = dCmPart_.as<double>();
    cp.cmCount = dCmCount_.as<unsigned int>();
    cp.cmDefer = 0;
    // batched form: the wide group CVs ride in the first coordination
    // item's rank launch (1,024-thread items) or in the path-row launch
    // (512-thread items, 8 phases) when the plan has one (no data between
    // them), else they take a launch of their own
    const GroupExtra ge{dGroupBig_.as<GroupCVItem>(), (int)groupBig_.size(), dGroupPos_.as<int>(),
                        dGroupMass_.as<float>(), dGroupRef_.as<float>()};
    bool bigPending = false;
    if (fused && w0 == 0 && nw == W_) {
        // a plan with a path CV: its row launch adds the contact-map partials
        cp.cmDefer = !path_.empty() && !cm_.empty() ? 1 : 0;
        launch_cv_fused(v, cp, nw, s);
        bigPending = !groupBig_.empty();
        const bool pathCarries = groupBigNT_ == 512 && !path_.empty() && tune_.pathPhases == 8;
        if (bigPending && coord_.empty() && !pathCarries) {
            launch_cv_group_big(v, ge.items, ge.n, ge.slots, ge.mass, ge.ref, w0, nw, groupBigNT_, s);
            bigPending = false;
        }
        GroupExtra px = ge;
        if (!(bigPending && pathCarries)) px.n = 0;
        if (cp.cmDefer) {
            px.cm = cp.cm;
            px.nCm = cp.nCm;
            px.nCmBlocks = cp.nCmBlocks;
            px.cmPart = cp.cmPart;
        }
        if (launch_path_rows(v, dPath_.as<PathItem>(), (int)path_.size(), pathRowBlocks_, pathMaxN_,
                             dPathPos_.as<int>(), dPathFrames_.as<float>(), w0, nw, tune_.pathPhases, s,
                             px.n > 0 || px.nCm > 0 ? &px : nullptr))
            bigPending = false;
    } else {
        launch_cv_small(v, dSmall_.as<SmallCVItem>(), (int)small_.size(), w0, nw, s);
        launch_cv_group(v, dGroup_.as<GroupCVItem>(), (int)group_.size(), dGroupPos_.as<int>(),
                        dGroupMass_.as<float>(), dGroupRef_.as<float>(), w0, nw, s);
        launch_cv_cm_chunks(v, cp, w0, nw, s);
        launch_cv_group_big(v, dGroupBig_.as<GroupCVItem>(), (int)groupBig_.size(), dGroupPos_.as<int>(),
                            dGroupMass_.as<float>(), dGroupRef_.as<float>(), w0, nw, groupBigNT_, s);
        launch_cv_path(v, dPath_.as<PathItem>(), (int)path_.size(), pathMsdUnits_, pathRowBlocks_, pathMaxN_,
                       dPathPos_.as<int>(), dPathFrames_.as<float>(), w0, nw, tune_.pathPhases, s);
    }
    if (!coord_.empty()) {
        CoordBuffers cb;
        cb.slots = dCoordSlots_.as<int>();
        cb.pos = dCoordPos_.as<int>();
        cb.posStride = std::max((int)coordSlots_.size(), 1);
        cb.cellOf = dCellOf_.as<int>();
        cb.cellCount = dCellCount_.as<int>();
        cb.cellStart = dCellStart_.as<int>();
        cb.sorted = dSorted_.as<float4>();
        cb.unsorted = dUnsorted_.as<float4>();
        cb.rank = dRank_.as<int>();
        cb.blockCnt = dBlockCnt_.as<int>();
        cb.blkStride = blkStride_;
        cb.cnAcc = dCnAcc_.as<long long>();
        cb.cnAccLo = dCnAccLo_.as<unsigned long long>();
        cb.pairAcc = dPairAcc_.as<unsigned long long>();
        cb.npairs = dNpairs_.as<long long>();
        cb.ticket = dCoordTicket_.as<unsigned int>();
        cb.cellOfStride = cellOfStride_;
        cb.cellStride = cellStride_;
        cb.sortStride = sortStride_;
        cb.nItems = (int)coord_.size();
        for (size_t i = 0; i < coord_.size(); i++)
            if (launch_coordination(v, cb, dCoord_.as<CoordItem>(), coord_[i], (int)i, w0, nw, s,
                                    bigPending && groupBigNT_ == 1024 ? &ge : nullptr))
                bigPending = false;
    }
    if (bigPending) launch_cv_group_big(v, ge.items, ge.n, ge.slots, ge.mass, ge.ref, w0, nw, groupBigNT_, s);
    }
    SmallBiasState st;
    st.grids = dGridTable_.as<GridBiasDev>();
    st.stateConst = dSmallConst_.as<double>();
    st.stateVar = dSmallVar_.as<double>();
    st.scal = dSmallScal_.as<double>();
    st.stateVarStride = smallVarPerWalker_;
    st.scalStride = 4 * (long long)std::max((int)smallBias_.size(), 1);
    st.pbGrid = dPbGrid_.as<double>();
    const ScatterPlan sp = scatterPlan();
    const bool biases = (stageMask & 10) && testGrads_.empty();
    if (biases || (stageMask & 4)) {
        for (auto* o : opes_)
            if (biases && o->d.adaptiveStride > 0 && (stageMask & 2)) {
                launch_opes_welford(v, o->d, w0, nw, s);
                if (!capturing_) bumpWelfordMirrors(o, w0, nw);
            }
    }
    bool scattered = false, denseDone = false;
    if (fused && (stageMask & 6) == 6 && w0 == 0 && nw == W_) {
        // batched: chunked sums + finalize + one-thread biases + dU/dCV (+ scatter) in one launch
        FusedBiasPlan fp = fplan_;
        fp.st = st;
        fp.sp = sp;
        if (!biases) {
            fp.nChunked = 0;
            fp.totalChunks = 0;
            fp.nSmall = 0;
        }
        fp.scatter = fplan_.scatter && forces;
        const bool dense = energyValue_ >= 0 && dSlotMask_.bytes;
        fp.finalizeScatter = forces && !fp.scatter && !dense;
        launch_step_bias(v, fp, nw, s);
        scattered = fp.scatter != 0 || fp.finalizeScatter != 0;
        if (forces && !scattered) {
            if (energyValue_ >= 0 && dSlotMask_.bytes) {
                // slots and the dense energy CV in one pass (dU/dCV from the finalize)
                launch_scatter_dense_rows(v, sp, dEnergyFPtr_.as<const float*>(), energyValue_, dSlotList_.as<int2>(),
                                          dSlotMask_.as<uint2>(), std::max(S_, 1), maskStride_, maxPadded_, w0, nw,
                                          s);
                denseDone = true;
            } else {
                launch_scatter(v, sp, w0, nw, 1, s);
            }
        }
        scattered = true;
    } else if (biases) {
        for (auto* o : opes_) launch_bias_opes(v, o->d, w0, nw, s);
        for (auto* h : hills_) launch_bias_hills(v, h->d, w0, nw, s);
        if (stageMask & 2)
            launch_bias_small(v, st, dSmallBias_.as<SmallBiasItem>(), (int)smallBias_.size(), w0, nw, s);
    }
    if (!(stageMask & 4)) {
        hip_check(hipGetLastError(), "kernel launch");
        return;
    }
    if (!scattered) launch_scatter(v, sp, w0, nw, forces ? 1 : 0, s);
    if (forces && energyValue_ >= 0 && !denseDone)
        launch_scatter_dense(v, dEnergyFPtr_.as<const float*>(), energyValue_, maxPadded_, w0, nw, s);
    hip_check(hipGetLastError(), "kernel launch");
}

ScatterPlan Engine::scatterPlan() const
{
    ScatterPlan sp;
    sp.segStart = dSegStart_.as<int>();
    sp.entryCv = dRowCv_.as<int>();
    sp.combStart = dCombStart_.as<int>();
    sp.combSlot = dCombSlot_.as<int>();
    sp.nComb = (int)combSlot_.size();
    sp.testGrad = testGrads_.empty() ? nullptr : dTestGrad_.as<double>();
    sp.energyValue = energyValue_;
    return sp;
}

// The host mirror of the adaptive-sigma sample count (the reference keeps
// the same mirror, CommonGluedKernels.cpp:4972-4976 and :5097-5102).
void Engine::bumpWelfordMirrors(OpesHost* o, int w0, int nw)
{
    for (int t = 0; t < (int)o->nSamplesMirror.size(); t++)
        if (o->d.shared || (t >= w0 && t < w0 + nw)) o->nSamplesMirror[t] += o->d.shared ? nw : 1;
}

void Engine::sync(hipStream_t s) { hip_check(hipStreamSynchronize(s), "hipStreamSynchronize"); }

// Plain form: one walker, one launch per CV kind and bias, then the scatter.
double Engine::execute(hipStream_t s, int w, const float4* posq, long long* force, bool forces, bool energy)
{
    if (!initialized_) fail("execute: initialize first");
    if (w < 0 || w >= W_) fail("execute: walker out of range");
    if (V_ == 0) return 0.0;
    // posq / force point at this walker's block; present them as the batch base
    lastPos_ = posq - posOff_[w];
    lastForce_ = force - forceOff_[w];
    if (!tuned_) {  // a one-walker first evaluation: no timing, a cached choice if there is one
        tuned_ = true;
        if (!evaluated_) adoptCachedTuning();
    }
    launchStep(s, w, 1, forces);
    cvReady_ = true;
    evaluated_ = true;
    double e = 0.0;
    if (energy && testGrads_.empty() && !biases_.empty()) {
        std::vector<double> be((size_t)B_);
        hip_check(hipMemcpyAsync(be.data(), dBener_.as<double>() + (size_t)w * B_, sizeof(double) * B_,
                                 hipMemcpyDeviceToHost, s),
                  "energy read-back");
        sync(s);
        for (size_t b = 0; b < biases_.size(); b++) e += be[b];
    }
    return e;
}

// The plain form for every walker, looped on the host side of the library:
// walker w's launch sequence is the one execute(w) issues, back to back on s.
void Engine::executePlainAll(hipStream_t s, const float4* posq, long long* force, bool forces)
{
    if (!initialized_) fail("execute_plain_all: initialize first");
    if (V_ == 0) return;
    lastPos_ = posq;
    lastForce_ = force;
    if (!tuned_) {
        tuned_ = true;
        if (!evaluated_) adoptCachedTuning();
    }
    for (int w = 0; w < W_; w++) launchStep(s, w, 1, forces);
    cvReady_ = true;
    evaluated_ = true;
}

void Engine::executeStage(hipStream_t s, const float4* posq, long long* force, int stage)
{
    if (!initialized_) fail("execute_stage: initialize first");
    if (stage < 0 || stage > 4) fail("execute_stage: stage must be 0, 1, 2, 3 or 4");
    lastPos_ = posq;
    lastForce_ = force;
    if (stage == 4) {
        // the force scatter of the batched step alone when it is the combined
        // dense + slot pass (energy CV, slots not scattered by the finalize),
        // else the dense energy-CV scatter alone; every walker, dU/dCV of the
        // last evaluation
        if (energyValue_ >= 0) {
            StepView v = view(lastPos_, lastForce_);
            v.S = S_;
            if (dSlotMask_.bytes && !fplan_.scatter)
                launch_scatter_dense_rows(v, scatterPlan(), dEnergyFPtr_.as<const float*>(), energyValue_,
                                          dSlotList_.as<int2>(), dSlotMask_.as<uint2>(), std::max(S_, 1), maskStride_,
                                          maxPadded_, 0, W_, s);
            else
                launch_scatter_dense(v, dEnergyFPtr_.as<const float*>(), energyValue_, maxPadded_, 0, W_, s);
        }
        hip_check(hipGetLastError(), "kernel launch");
        return;
    }
    if (stage == 3) {
        // the batched chunk kernel alone (every hill / OPES chunk of every walker)
        if (fplan_.totalChunks > 0 && testGrads_.empty()) {
            StepView v = view(lastPos_, lastForce_);
            v.S = S_;
            launch_step_bias_chunks(v, fplan_, W_, s);
        }
        hip_check(hipGetLastError(), "kernel launch");
        return;
    }
    if (!tuned_) {
        tuned_ = true;
        if (!evaluated_) adoptCachedTuning();
    }
    evaluated_ = true;
    launchStep(s, 0, W_, true, 1 << stage);
}

// Batched form: every walker in one launch per stage, optionally as one
// captured hipGraph (re-captured when the buffers change).
void Engine::executeBatch(hipStream_t s, const float4* posq, long long* force, bool forces, bool useGraph)
{
    if (!initialized_) fail("execute_batch: initialize first");
    if (V_ == 0) return;
    lastPos_ = posq;
    lastForce_ = force;
    autotune(s);
    evaluated_ = true;
    if (!useGraph) {
        launchStep(s, 0, W_, forces, 7, true);
        cvReady_ = true;
        return;
    }
    if (graphPos_ != posq || graphForce_ != force || graphForces_ != forces || !graphExec_) {
        if (graphExec_) {
            (void)hipGraphExecDestroy(graphExec_);
            (void)hipGraphDestroy(graph_);
            graphExec_ = nullptr;
            graph_ = nullptr;
        }
        // Capture on a private stream (the caller's may be the legacy default
        // stream, which cannot be captured); the graph is launched on the
        // caller's stream, so ordering with the caller's work is preserved.
        if (!captureStream_) hip_check(hipStreamCreateWithFlags(&captureStream_, hipStreamNonBlocking), "stream");
        hip_check(hipStreamBeginCapture(captureStream_, hipStreamCaptureModeThreadLocal), "begin capture");
        capturing_ = true;
        try {
            launchStep(captureStream_, 0, W_, forces, 7, true);
            capturing_ = false;
        } catch (...) {
            capturing_ = false;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(captureStream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(captureStream_, &graph_), "end capture");
        hip_check(hipGraphInstantiate(&graphExec_, graph_, nullptr, nullptr, 0), "graph instantiate");
        graphPos_ = posq;
        graphForce_ = force;
        graphForces_ = forces;
    }
    hip_check(hipGraphLaunch(graphExec_, s), "graph launch");
    if (testGrads_.empty())
        for (auto* o : opes_)
            if (o->d.adaptiveStride > 0) bumpWelfordMirrors(o, 0, W_);
    cvReady_ = true;
}

// ---------------------------------------------------------------------------
// Deposition, CommonGluedKernels.cpp:5050-5157 + GluedForceImpl.cpp:100-109:
// skipped until one evaluation filled the CVs, once per step, never at step 0,
// pace-gated; OPES, then OPES-expanded, multithermal, then MetaD.
// ---------------------------------------------------------------------------
int Engine::depositSites() const
{
    return (int)(opes_.size() + smallBias_.size() + grids_.size() + hills_.size());
}

// The gating of updateState (CommonGluedKernels.cpp:5050-5157): nothing at
// step 0; each site at its own pace; OPES only while its table has room.
uint64_t Engine::depositMask(int step) const
{
    uint64_t m = 0;
    int bit = 0;
    if (step == 0) return 0;
    for (auto* o : opes_) {
        bool any = false;
        for (int c : o->depositMirror) any |= c < o->d.cap;
        if (step % o->pace == 0 && any) m |= 1ull << bit;
        bit++;
    }
    for (size_t i = 0; i < smallBias_.size(); i++) {
        const auto& hs = smallBiasHost_[i];
        const bool dep = hs.type == BIAS_OPES_EXPANDED ? step % hs.pace == 0
                         : hs.type == BIAS_OPES_MULTITHERMAL ? (hs.pace > 0 && step % hs.pace == 0)
                         : hs.type == BIAS_PBMETAD           ? step % hs.pace == 0
                                                             : false;
        if (dep) m |= 1ull << bit;
        bit++;
    }
    for (auto* g : grids_) {
        if (step % g->pace == 0) m |= 1ull << bit;
        bit++;
    }
    for (auto* h : hills_) {
        if (step % h->pace == 0) m |= 1ull << bit;
        bit++;
    }
    return m;
}

// Deposit launches in the reference's order: OPES, OPES-expanded,
// multithermal, grid MetaD, hill lists.
void Engine::launchDeposits(hipStream_t s, uint64_t mask)
{
    StepView v = view(lastPos_, lastForce_);
    v.S = S_;
    const int nO = (int)opes_.size(), nS = (int)smallBias_.size(), nG = (int)grids_.size();
    for (int k = 0; k < nO; k++)
        if (mask >> k & 1) launch_dep_opes(v, opes_[k]->d, 0, W_, s);
    SmallBiasState st;
    st.grids = dGridTable_.as<GridBiasDev>();
    st.stateConst = dSmallConst_.as<double>();
    st.stateVar = dSmallVar_.as<double>();
    st.scal = dSmallScal_.as<double>();
    st.stateVarStride = smallVarPerWalker_;
    st.scalStride = 4 * (long long)std::max(nS, 1);
    st.pbGrid = dPbGrid_.as<double>();
    for (int pass = 0; pass < 2; pass++)  // expanded first, then multithermal
        for (int i = 0; i < nS; i++) {
            const int want = pass == 0 ? BIAS_OPES_EXPANDED : BIAS_OPES_MULTITHERMAL;
            if (smallBiasHost_[i].type == want && (mask >> (nO + i) & 1))
                launch_dep_small(v, st, dSmallBias_.as<SmallBiasItem>(), nS, i, 0, W_, s);
        }
    for (int k = 0; k < nG; k++)
        if (mask >> (nO + nS + k) & 1) launch_dep_grid(v, grids_[k]->d, 0, W_, s);
    for (auto& ph : pbs_)  // after the MetaD grids (CommonGluedKernels.cpp:5147-5154)
        if (mask >> (nO + ph.item) & 1) launch_dep_pb(v, st, smallBias_[ph.item], 0, W_, s);
    for (int k = 0; k < (int)hills_.size(); k++)
        if (mask >> (nO + nS + nG + k) & 1) launch_dep_hills(v, hills_[k]->d, 0, W_, s);
    hip_check(hipGetLastError(), "deposit launch");
}

void Engine::commitDeposits(uint64_t mask)
{
    const int nO = (int)opes_.size(), nS = (int)smallBias_.size();
    for (int k = 0; k < nO; k++) {
        if (!(mask >> k & 1)) continue;
        OpesHost* o = opes_[k];
        for (size_t t = 0; t < o->depositMirror.size(); t++) {
            const bool will = o->d.adaptiveStride == 0 || o->nSamplesMirror[t] >= o->d.adaptiveStride;
            if (will) o->depositMirror[t] += o->d.shared ? W_ : 1;
        }
    }
    for (size_t k = 0; k < grids_.size(); k++)
        if (mask >> (nO + nS + k) & 1)
            for (auto& n : grids_[k]->numDeposited) n += grids_[k]->d.walkerStride ? 1 : W_;
    for (auto& ph : pbs_)
        if (mask >> (nO + ph.item) & 1)
            for (auto& n : ph.numDeposited) n += ph.T == 1 ? W_ : 1;
}

void Engine::updateState(hipStream_t s, int step)
{
    if (!initialized_) fail("update_state: initialize first");
    if (!cvReady_) return;
    if (step == lastUpdateStep_) return;
    if (depositSites() == 0) return;
    if (depositSites() > 64) fail("update_state: more than 64 deposit sites");
    lastUpdateStep_ = step;
    const uint64_t m = depositMask(step);
    launchDeposits(s, m);
    commitDeposits(m);
}

void Engine::dropGraphs()
{
    if (graphExec_) (void)hipGraphExecDestroy(graphExec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graphExec_ = nullptr;
    graph_ = nullptr;
    for (auto& kv : stepGraphs_) {
        (void)hipGraphExecDestroy(kv.second.second);
        (void)hipGraphDestroy(kv.second.first);
    }
    stepGraphs_.clear();
}

// updateState(step) followed by the batched evaluation (the order OpenMM
// uses: updateContextState, then calcForcesAndEnergy), as one graph launch.
// Graphs are cached per deposit mask; the host mirrors advance exactly as
// updateState + executeBatch would advance them.
void Engine::stepBatch(hipStream_t s, int step, const float4* posq, long long* force, bool forces)
{
    if (!initialized_) fail("step_batch: initialize first");
    if (V_ == 0) return;
    if (depositSites() > 64) fail("step_batch: more than 64 deposit sites");
    uint64_t m = 0;
    if (cvReady_ && step != lastUpdateStep_ && depositSites() > 0) {
        m = depositMask(step);
        lastUpdateStep_ = step;
    }
    lastPos_ = posq;
    lastForce_ = force;
    autotune(s);
    evaluated_ = true;
    if (stepGraphPos_ != posq || stepGraphForce_ != force || stepGraphForces_ != forces) {
        for (auto& kv : stepGraphs_) {
            (void)hipGraphExecDestroy(kv.second.second);
            (void)hipGraphDestroy(kv.second.first);
        }
        stepGraphs_.clear();
        stepGraphPos_ = posq;
        stepGraphForce_ = force;
        stepGraphForces_ = forces;
    }
    auto it = stepGraphs_.find(m);
    if (it == stepGraphs_.end()) {
        if (!captureStream_) hip_check(hipStreamCreateWithFlags(&captureStream_, hipStreamNonBlocking), "stream");
        hip_check(hipStreamBeginCapture(captureStream_, hipStreamCaptureModeThreadLocal), "begin capture");
        capturing_ = true;
        try {
            if (m) launchDeposits(captureStream_, m);
            launchStep(captureStream_, 0, W_, forces, 7, true);
            capturing_ = false;
        } catch (...) {
            capturing_ = false;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(captureStream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        hip_check(hipStreamEndCapture(captureStream_, &g), "end capture");
        hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "graph instantiate");
        it = stepGraphs_.emplace(m, std::make_pair(g, ge)).first;
    }
    hip_check(hipGraphLaunch(it->second.second, s), "graph launch");
    if (m) commitDeposits(m);
    if (testGrads_.empty())
        for (auto* o : opes_)
            if (o->d.adaptiveStride > 0) bumpWelfordMirrors(o, 0, W_);
    cvReady_ = true;
}

// ---------------------------------------------------------------------------
// Autograd support: device-to-device state copies, differentiable
// parameters, and the hill adjoint.
// ---------------------------------------------------------------------------
void Engine::copyBuffer(int which, void* dst, size_t bytes, hipStream_t s)
{
    if (!initialized_) fail("copy_buffer: initialize first");
    const void* src = nullptr;
    size_t n = 0;
    switch (which) {
        case 0: src = dCv_.p; n = sizeof(double) * (size_t)W_ * std::max(V_, 1); break;
        case 1: src = dBgrad_.p; n = sizeof(double) * (size_t)W_ * G_; break;
        case 2: src = dBener_.p; n = sizeof(double) * (size_t)W_ * B_; break;
        case 3: src = dUd_.p; n = sizeof(double) * (size_t)W_ * std::max(V_, 1); break;
        case 4: src = dPathAux_.p; n = sizeof(double) * (size_t)W_ * std::max((int)path_.size(), 1) * 2; break;
        default: fail("copy_buffer: unknown buffer id");
    }
    if (bytes < n) fail("copy_buffer: destination smaller than the buffer (" + std::to_string(n) + " bytes)");
    hip_check(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, s), "copy_buffer");
}

void Engine::setHillHeightsDevice(int w, int bi, const double* d, int n, hipStream_t s)
{
    HillHost* h = nullptr;
    for (auto* x : hills_)
        if (x->spec == bi) h = x;
    if (!h) fail("set_hill_heights: not a hill bias");
    if (w < 0 || w >= W_) fail("set_hill_heights: walker out of range");
    if (n < 0 || n > h->d.cap) fail("set_hill_heights: count exceeds the capacity");
    const int t = h->d.shared ? 0 : w;
    launch_f64_to_f32(d, h->d.heights + (size_t)t * h->d.cap, n, s);
    hip_check(hipGetLastError(), "set_hill_heights");
}

void Engine::setPathLambda(int pi, double lambda)
{
    if (!initialized_) fail("set_path_lambda: initialize first");
    if (pi < 0 || pi >= (int)path_.size()) fail("set_path_lambda: path index out of range");
    if (!(lambda > 0.0)) fail("set_path_lambda: lambda must be positive");
    path_[pi].lambda = lambda;
    hip_check(hipMemcpy(dPath_.as<PathItem>() + pi, &path_[pi], sizeof(PathItem), hipMemcpyHostToDevice),
              "set_path_lambda");
}

void Engine::setHillSigma(int bi, const double* sigma, int D)
{
    HillHost* h = nullptr;
    int k = 0;
    for (size_t i = 0; i < hills_.size(); i++)
        if (hills_[i]->spec == bi) {
            h = hills_[i];
            k = (int)i;
        }
    if (!h) fail("set_hill_sigma: not a hill bias");
    if (D != h->d.D) fail("set_hill_sigma: dimension mismatch");
    for (int d = 0; d < D; d++) {
        if (!(sigma[d] > 0.0)) fail("set_hill_sigma: sigma must be positive");
        h->d.sigma[d] = sigma[d];
        h->d.i2s2[d] = 1.0 / (2.0 * sigma[d] * sigma[d]);
        h->d.is2[d] = 1.0 / (sigma[d] * sigma[d]);
    }
    // the batched stage reads the table from device memory, and the first
    // hill bias also from its copy in the launch arguments (captured graphs
    // hold that copy: they are rebuilt)
    hip_check(hipMemcpy(dHillTable_.as<HillBiasDev>() + k, &h->d, sizeof(HillBiasDev), hipMemcpyHostToDevice),
              "set_hill_sigma");
    if (k == 0) {
        fplan_.hill0 = h->d;
        dropGraphs();
    }
}

size_t Engine::hillAdjointPartials(int bi) const
{
    for (auto* x : hills_)
        if (x->spec == bi) return (size_t)hill_adjoint_blocks(x->d.cap) * (x->d.shared ? 1 : W_) * x->d.D;
    fail("hill_adjoint: not a hill bias");
    return 0;
}

void Engine::hillAdjoint(hipStream_t s, int bi, const double* gV, double* gH, double* gSigma, double* part,
                         size_t partCap)
{
    if (!initialized_ || !cvReady_) fail("hill_adjoint: evaluate first");
    HillHost* h = nullptr;
    for (auto* x : hills_)
        if (x->spec == bi) h = x;
    if (!h) fail("hill_adjoint: not a hill bias");
    const size_t need = hillAdjointPartials(bi);
    if (part) {
        if (need > partCap)
            fail("hill_adjoint: needs " + std::to_string(need) + " doubles of partials, the buffer holds " +
                 std::to_string(partCap));
    } else {
        if (dAdjPart_.bytes < sizeof(double) * need) dAdjPart_.alloc(sizeof(double) * need);
        part = dAdjPart_.as<double>();
    }
    StepView v = view(lastPos_, lastForce_);
    v.S = S_;
    launch_hill_adjoint(v, h->d, gV, gH, part, gSigma, W_, s);
    hip_check(hipGetLastError(), "hill_adjoint launch");
}

// Frames as the CV stores them (FP32, [N][M][3]); the spec keeps the new
// values so that a re-plan uses them too.
void Engine::setPathFrames(int pi, const double* frames, int n)
{
    if (!initialized_) fail("set_path_frames: initialize first");
    if (pi < 0 || pi >= (int)path_.size()) fail("set_path_frames: path index out of range");
    const PathItem& it = path_[pi];
    if (n != it.N * it.M * 3) fail("set_path_frames: expected N*M*3 = " + std::to_string(it.N * it.M * 3) + " values");
    std::vector<float> f((size_t)n);
    for (int k = 0; k < n; k++) {
        if (!std::isfinite(frames[k])) fail("set_path_frames: non-finite coordinate");
        f[k] = (float)frames[k];
    }
    int seen = 0;
    for (auto& c : cvs_)
        if (c.type == CV_PATH && seen++ == pi)
            for (int k = 0; k < n; k++) c.params[2 + k] = (double)f[k];
    hip_check(hipMemcpy(dPathFrames_.as<float>() + it.frameOff, f.data(), sizeof(float) * n, hipMemcpyHostToDevice),
              "set_path_frames");
    centrePath(pi, f.data());
    hip_check(hipMemcpy(dPathCentre_.as<float>() + it.centreOff, pathCentreHost_.data() + it.centreOff,
                        sizeof(float) * (size_t)3 * it.M, hipMemcpyHostToDevice),
              "set_path_frames (centre)");
    hip_check(hipMemcpy(dPathR2_.as<double>() + it.msdOff, pathR2Host_.data() + it.msdOff, sizeof(double) * it.N,
                        hipMemcpyHostToDevice),
              "set_path_frames (|r'|^2)");
}

// Centre of path pi (the mean frame, rounded to FP32) and |r_f - c|^2 in
// FP64 (r - c is exact in FP64 for FP32 r and c, as in the MSD blocks).
void Engine::centrePath(int pi, const float* frames)
{
    const PathItem& it = path_[pi];
    const int M = it.M, N = it.N, K = 3 * M;
    float* c = pathCentreHost_.data() + it.centreOff;
    for (int k = 0; k < K; k++) {
        double m = 0.0;
        for (int f = 0; f < N; f++) m += (double)frames[(size_t)f * K + k];
        c[k] = (float)(m / N);
    }
    for (int f = 0; f < N; f++) {
        double r2 = 0.0;
        for (int k = 0; k < K; k++) {
            const double d = (double)frames[(size_t)f * K + k] - (double)c[k];
            r2 += d * d;
        }
        pathR2Host_[it.msdOff + f] = r2;
    }
}

void Engine::pathFramesAdjoint(hipStream_t s, int pi, const double* gV, double* gF)
{
    if (!initialized_ || !cvReady_) fail("path_frames_adjoint: evaluate first");
    if (pi < 0 || pi >= (int)path_.size()) fail("path_frames_adjoint: path index out of range");
    const PathItem& it = path_[pi];
    const size_t need = sizeof(double) * (size_t)W_ * it.N;
    if (dPathCoef_.bytes < need) dPathCoef_.alloc(need);
    StepView v = view(lastPos_, lastForce_);
    v.S = S_;
    launch_path_frames_adjoint(v, dPath_.as<PathItem>(), pi, it.N, it.M, dPathPos_.as<int>(),
                               dPathFrames_.as<float>(), gV, dPathCoef_.as<double>(), gF, W_, s);
    hip_check(hipGetLastError(), "path_frames_adjoint launch");
}

// ---------------------------------------------------------------------------
// Read-backs
// ---------------------------------------------------------------------------
template <class T>
static std::vector<T> download(const T* p, size_t n)
{
    std::vector<T> h(n);
    if (n) hip_check(hipMemcpy(h.data(), p, sizeof(T) * n, hipMemcpyDeviceToHost), "download");
    return h;
}

std::vector<double> Engine::cvValues(int w)
{
    if (!initialized_ || w < 0 || w >= W_) fail("get_cv_values: bad walker or not initialized");
    hip_check(hipDeviceSynchronize(), "sync");
    return download(dCv_.as<double>() + (size_t)w * std::max(V_, 1), V_);
}

std::vector<double> Engine::biasEnergies(int w)
{
    if (!initialized_ || w < 0 || w >= W_) fail("get_bias_energies: bad walker or not initialized");
    hip_check(hipDeviceSynchronize(), "sync");
    return download(dBener_.as<double>() + (size_t)w * B_, biases_.size());
}

std::vector<double> Engine::dUdCV(int w)
{
    if (!initialized_ || w < 0 || w >= W_) fail("get_dudcv: bad walker or not initialized");
    hip_check(hipDeviceSynchronize(), "sync");
    return download(dUd_.as<double>() + (size_t)w * std::max(V_, 1), V_);
}

long long Engine::coordinationPairs(int w, int ci)
{
    if (!initialized_ || w < 0 || w >= W_) fail("coordination_pairs: bad walker or not initialized");
    if (ci < 0 || ci >= (int)coord_.size()) fail("coordination_pairs: index out of range");
    hip_check(hipDeviceSynchronize(), "sync");
    return download(dNpairs_.as<long long>() + (size_t)w * coord_.size() + ci, 1)[0];
}

// Adjoint of an OPES bias with respect to its kernel table (centres,
// widths, log-weights) and sum_uprob, at the CV values and energies of the
// last evaluation; g = dL/dV_w on the device, [W].
void Engine::opesAdjoint(hipStream_t s, int bi, const double* gV, double* gC, double* gS, double* gL, double* gU)
{
    if (!initialized_ || !cvReady_) fail("opes_adjoint: evaluate first");
    OpesHost* o = nullptr;
    for (auto* x : opes_)
        if (x->spec == bi) o = x;
    if (!o) fail("opes_adjoint: not an OPES bias");
    StepView v = view(lastPos_, lastForce_);
    v.S = S_;
    launch_opes_adjoint(v, o->d, gV, gC, gS, gL, gU, W_, s);
    hip_check(hipGetLastError(), "opes_adjoint launch");
}

// A kernel table set directly (kernel-major centres and widths [n][D],
// log-weights [n]) with its sum_uprob; the count and the slot-claim counter
// follow n, the other accumulators are kept.
void Engine::setOpesKernels(int w, int bi, const double* centers, const double* sigmas, const double* logw, int n,
                            double sumUprob)
{
    if (!initialized_) fail("set_opes_kernels: initialize first");
    OpesHost* o = nullptr;
    for (auto* x : opes_)
        if (x->spec == bi) o = x;
    if (!o) fail("set_opes_kernels: not an OPES bias");
    if (w < 0 || w >= W_) fail("set_opes_kernels: walker out of range");
    if (n < 0 || n > o->d.cap) fail("set_opes_kernels: kernel count exceeds the capacity");
    hip_check(hipDeviceSynchronize(), "sync");  // no evaluation in flight reads the table
    const int t = o->d.shared ? 0 : w, D = o->d.D;
    std::vector<double> cs((size_t)D * o->d.cap, 0.0), ss((size_t)D * o->d.cap, 1.0);
    for (int k = 0; k < n; k++)
        for (int d = 0; d < D; d++) {
            cs[(size_t)d * o->d.cap + k] = centers[(size_t)k * D + d];
            ss[(size_t)d * o->d.cap + k] = sigmas[(size_t)k * D + d];
        }
    hip_check(hipMemcpy(o->d.centers + (size_t)t * D * o->d.cap, cs.data(), 8 * cs.size(), hipMemcpyHostToDevice),
              "set_opes_kernels");
    hip_check(hipMemcpy(o->d.sigmas + (size_t)t * D * o->d.cap, ss.data(), 8 * ss.size(), hipMemcpyHostToDevice),
              "set_opes_kernels");
    if (n) hip_check(hipMemcpy(o->d.logw + (size_t)t * o->d.cap, logw, 8 * (size_t)n, hipMemcpyHostToDevice),
                     "set_opes_kernels");
    auto I = download(o->d.ints + 4 * t, 4);
    I[0] = n;
    I[1] = n;
    hip_check(hipMemcpy(o->d.ints + 4 * t, I.data(), 16, hipMemcpyHostToDevice), "set_opes_kernels");
    hip_check(hipMemcpy(o->d.dbl + (size_t)t * (4 + 2 * D), &sumUprob, 8, hipMemcpyHostToDevice), "set_opes_kernels");
    o->depositMirror[t] = n;
}

// getOPESMetrics, CommonGluedKernels.cpp:5164-5210: [zed, rct, nker, neff].
std::vector<double> Engine::opesMetrics(int w, int oi)
{
    if (oi < 0 || oi >= (int)opes_.size()) fail("getOPESMetrics: biasIndex out of range");
    if (w < 0 || w >= W_) fail("getOPESMetrics: walker out of range");
    OpesHost* o = opes_[oi];
    const int t = o->d.shared ? 0 : w, D = o->d.D;
    hip_check(hipDeviceSynchronize(), "sync");
    auto I = download(o->d.ints + 4 * t, 4);
    auto Q = download(o->d.dbl + (size_t)t * (4 + 2 * D), 4);
    const int nker = I[0];
    const double counter = I[2];
    const double KDEnorm = o->d.variant == 2 ? counter : Q[1];
    const double zed = (nker > 0 && KDEnorm > 0.0) ? Q[0] / (KDEnorm * nker) : 1.0;
    const double rct = (counter > 0.0 && Q[1] > 0.0) ? o->d.kT * std::log(Q[1] / counter) : 0.0;
    const double neff = (1.0 + Q[2]) * (1.0 + Q[2]) / (1.0 + Q[3]);
    return {zed, rct, (double)nker, neff};
}

// getKernelSigmas, CommonGluedKernels.cpp:5212-5232 (kernel-major, CV inner).
std::vector<float> Engine::kernelSigmas(int w, int oi)
{
    if (oi < 0 || oi >= (int)opes_.size()) return {};
    OpesHost* o = opes_[oi];
    const int t = o->d.shared ? 0 : w, D = o->d.D;
    hip_check(hipDeviceSynchronize(), "sync");
    const int nk = download(o->d.ints + 4 * t, 1)[0];
    if (nk <= 0) return {};
    auto sg = download(o->d.sigmas + (size_t)t * D * o->d.cap, (size_t)D * o->d.cap);
    std::vector<float> out((size_t)nk * D);
    for (int k = 0; k < nk; k++)
        for (int d = 0; d < D; d++) out[(size_t)k * D + d] = (float)sg[(size_t)d * o->d.cap + k];
    return out;
}

void Engine::setHills(int w, int bi, const double* centers, const double* heights, int n)
{
    if (!initialized_) fail("set_hills: initialize first");
    if (bi < 0 || bi >= (int)biases_.size() || biases_[bi].type != BIAS_METAD_HILLS) fail("set_hills: not a hill bias");
    HillHost* h = nullptr;
    for (auto* x : hills_)
        if (x->spec == bi) h = x;
    if (w < 0 || w >= W_) fail("set_hills: walker out of range");
    if (n < 0 || n > h->d.cap) fail("set_hills: hill count exceeds the capacity");
    const int t = h->d.shared ? 0 : w, D = h->d.D;
    // stored in FP32 (rounded to nearest), kernel-major -> SoA
    std::vector<float> c((size_t)D * h->d.cap, 0.0f), hh((size_t)n);
    for (int k = 0; k < n; k++) {
        for (int d = 0; d < D; d++) c[(size_t)d * h->d.cap + k] = (float)centers[(size_t)k * D + d];
        hh[k] = (float)heights[k];
    }
    hip_check(hipMemcpy(h->d.centers + (size_t)t * D * h->d.cap, c.data(), sizeof(float) * c.size(),
                        hipMemcpyHostToDevice),
              "set_hills");
    if (n) hip_check(hipMemcpy(h->d.heights + (size_t)t * h->d.cap, hh.data(), sizeof(float) * n, hipMemcpyHostToDevice),
                     "set_hills");
    hip_check(hipMemcpy(h->d.count + t, &n, sizeof(int), hipMemcpyHostToDevice), "set_hills");
}

int Engine::getHills(int w, int bi, double* centers, double* heights, int cap)
{
    HillHost* h = nullptr;
    for (auto* x : hills_)
        if (x->spec == bi) h = x;
    if (!h) fail("get_hills: not a hill bias");
    if (w < 0 || w >= W_) fail("get_hills: walker out of range");
    const int t = h->d.shared ? 0 : w, D = h->d.D;
    hip_check(hipDeviceSynchronize(), "sync");
    const int n = download(h->d.count + t, 1)[0];
    if (centers && heights) {
        if (cap < n) fail("get_hills: buffer too small");
        auto c = download(h->d.centers + (size_t)t * D * h->d.cap, (size_t)D * h->d.cap);
        auto hh = download(h->d.heights + (size_t)t * h->d.cap, n);
        for (int k = 0; k < n; k++) {
            for (int d = 0; d < D; d++) centers[(size_t)k * D + d] = (double)c[(size_t)d * h->d.cap + k];
            heights[k] = (double)hh[k];
        }
    }
    return n;
}

void Engine::setGrid(int w, int bi, const double* grid, int n)
{
    GridHost* g = nullptr;
    for (auto* x : grids_)
        if (x->spec == bi) g = x;
    if (!g) fail("set_grid: not a grid MetaD bias");
    if (w < 0 || w >= W_) fail("set_grid: walker out of range");
    if (n != g->d.total) fail("set_grid: size mismatch");
    const int t = g->d.walkerStride ? w : 0;
    hip_check(hipMemcpy(g->d.grid + (size_t)t * g->d.total, grid, sizeof(double) * n, hipMemcpyHostToDevice),
              "set_grid");
}

// ---------------------------------------------------------------------------
// Bias-state blob, 'GPUS' v1 (CommonSerialization.cpp:20-178 / 180-410):
// OPES, ABMD(0), MetaD grids, PBMetaD(0), external(0), linear(0), walls(0),
// OPES-expanded, ext-Lagrangian(0), EDS(0), MaxEnt(0), multithermal; then a
// trailing hill-list section (extension; a reference reader stops before it).
// ---------------------------------------------------------------------------
std::vector<char> Engine::getState(int w)
{
    if (!initialized_ || w < 0 || w >= W_) fail("getBiasState: bad walker or not initialized");
    hip_check(hipDeviceSynchronize(), "sync");
    std::vector<char> buf;
    auto put = [&](const void* p, size_t n) { buf.insert(buf.end(), (const char*)p, (const char*)p + n); };
    auto i32 = [&](int32_t x) { put(&x, 4); };
    auto f64 = [&](double x) { put(&x, 8); };
    put("GPUS", 4);
    i32(1);
    i32((int32_t)opes_.size());
    for (auto* o : opes_) {
        const int t = o->d.shared ? 0 : w, D = o->d.D;
        auto I = download(o->d.ints + 4 * t, 4);
        auto Q = download(o->d.dbl + (size_t)t * (4 + 2 * D), 4 + 2 * D);
        const int nk = I[0];
        auto c = download(o->d.centers + (size_t)t * D * o->d.cap, (size_t)D * o->d.cap);
        auto sg = download(o->d.sigmas + (size_t)t * D * o->d.cap, (size_t)D * o->d.cap);
        auto lw = download(o->d.logw + (size_t)t * o->d.cap, nk);
        i32(nk);
        f64(Q[0]);
        i32(I[3]);
        put(&Q[4], 8 * D);
        put(&Q[4 + D], 8 * D);
        for (int k = 0; k < nk; k++)
            for (int d = 0; d < D; d++) f64(c[(size_t)d * o->d.cap + k]);
        for (int k = 0; k < nk; k++)
            for (int d = 0; d < D; d++) f64(sg[(size_t)d * o->d.cap + k]);
        put(lw.data(), 8 * (size_t)nk);
    }
    i32(0);  // ABMD
    i32((int32_t)grids_.size());
    for (auto* g : grids_) {
        const int t = g->d.walkerStride ? w : 0;
        i32(g->numDeposited[t]);
        auto gr = download(g->d.grid + (size_t)t * g->d.total, g->d.total);
        put(gr.data(), 8 * gr.size());
    }
    i32((int32_t)pbs_.size());  // PBMetaD: per bias its sub-grids, each {numDeposited, values}
    for (auto& ph : pbs_) {
        const int t = ph.T == 1 ? 0 : w;
        i32(ph.N);
        auto gr = download(dPbGrid_.as<double>() + ph.base + (size_t)t * ph.stride, (size_t)ph.stride);
        size_t off = 0;
        for (int i = 0; i < ph.N; i++) {
            i32(ph.numDeposited[t]);
            put(gr.data() + off, 8 * (size_t)ph.npts[i]);
            off += ph.npts[i];
        }
    }
    i32(0);  // external
    {
        // linear and wall restraints are stateless: their counts only
        // (CommonSerialization.cpp:104-108)
        int nlin = 0, nwall = 0;
        for (auto& it : smallBias_) {
            nlin += it.type == BIAS_LINEAR;
            nwall += it.type == BIAS_UPPER_WALL || it.type == BIAS_LOWER_WALL;
        }
        i32(nlin);
        i32(nwall);
    }
    int noe = 0, nmt = 0;
    for (auto& it : smallBias_) {
        noe += it.type == BIAS_OPES_EXPANDED;
        nmt += it.type == BIAS_OPES_MULTITHERMAL;
    }
    const size_t nsb = std::max((int)smallBias_.size(), 1);
    auto scal = download(dSmallScal_.as<double>() + (size_t)w * nsb * 4, nsb * 4);
    i32(noe);
    for (size_t i = 0; i < smallBias_.size(); i++)
        if (smallBias_[i].type == BIAS_OPES_EXPANDED) {
            f64(scal[4 * i]);
            i32((int32_t)scal[4 * i + 1]);
        }
    i32(0);  // ext-Lagrangian
    i32(0);  // EDS
    i32(0);  // MaxEnt
    i32(nmt);
    auto var = download(dSmallVar_.as<double>() + (size_t)w * smallVarPerWalker_, smallVarPerWalker_);
    for (size_t i = 0; i < smallBias_.size(); i++)
        if (smallBias_[i].type == BIAS_OPES_MULTITHERMAL) {
            const auto& it = smallBias_[i];
            i32(it.nState);
            put(&var[it.stateOff], 8 * (size_t)it.nState);
            f64(scal[4 * i + 2]);
            f64(scal[4 * i + 3]);
        }
    // extension: hill lists
    put("HILL", 4);
    i32((int32_t)hills_.size());
    for (auto* h : hills_) {
        const int t = h->d.shared ? 0 : w, D = h->d.D;
        const int n = download(h->d.count + t, 1)[0];
        auto c = download(h->d.centers + (size_t)t * D * h->d.cap, (size_t)D * h->d.cap);
        auto hh = download(h->d.heights + (size_t)t * h->d.cap, n);
        i32(n);
        i32(D);
        for (int k = 0; k < n; k++)
            for (int d = 0; d < D; d++) f64((double)c[(size_t)d * h->d.cap + k]);
        for (int k = 0; k < n; k++) f64((double)hh[k]);
    }
    return buf;
}

void Engine::setState(int w, const std::vector<char>& blob)
{
    if (!initialized_ || w < 0 || w >= W_) fail("setBiasState: bad walker or not initialized");
    if (blob.empty()) return;
    const char* p = blob.data();
    const char* const end = p + blob.size();
    auto get = [&](void* dst, size_t n) {
        if ((size_t)(end - p) < n) fail("GLUED bias-state: truncated/corrupt blob");
        std::memcpy(dst, p, n);
        p += n;
    };
    auto i32 = [&]() {
        int32_t x;
        get(&x, 4);
        return x;
    };
    auto f64 = [&]() {
        double x;
        get(&x, 8);
        return x;
    };
    if ((size_t)(end - p) >= 8 && std::memcmp(p, "GPUS", 4) == 0) {
        p += 4;
        if (i32() != 1) fail("GLUED bias-state: unknown blob version");
    }
    auto count = [&](const char* what, int expect) {
        int32_t n = i32();
        if (n != expect)
            fail(std::string("GLUED bias-state: ") + what + " bias count mismatch (expected " + std::to_string(expect) +
                 ", got " + std::to_string(n) + ")");
    };
    count("OPES", (int)opes_.size());
    for (auto* o : opes_) {
        const int t = o->d.shared ? 0 : w, D = o->d.D;
        const int nk = i32();
        const double sumUprob = f64();
        const int ns = i32();
        if (nk < 0 || nk > o->d.cap) fail("GLUED bias-state: OPES kernel count out of range");
        const size_t need = (size_t)D * 16 + (size_t)nk * D * 16 + (size_t)nk * 8;
        if ((size_t)(end - p) < need) fail("GLUED bias-state: truncated/corrupt blob");
        std::vector<double> mean(D), M2(D), c((size_t)nk * D), sg((size_t)nk * D), lw(nk);
        get(mean.data(), 8 * D);
        get(M2.data(), 8 * D);
        get(c.data(), 8 * c.size());
        get(sg.data(), 8 * sg.size());
        get(lw.data(), 8 * lw.size());
        std::vector<double> cs((size_t)D * o->d.cap, 0.0), ss((size_t)D * o->d.cap, 1.0);
        for (int k = 0; k < nk; k++)
            for (int d = 0; d < D; d++) {
                cs[(size_t)d * o->d.cap + k] = c[(size_t)k * D + d];
                ss[(size_t)d * o->d.cap + k] = sg[(size_t)k * D + d];
            }
        hip_check(hipMemcpy(o->d.centers + (size_t)t * D * o->d.cap, cs.data(), 8 * cs.size(), hipMemcpyHostToDevice),
                  "restore");
        hip_check(hipMemcpy(o->d.sigmas + (size_t)t * D * o->d.cap, ss.data(), 8 * ss.size(), hipMemcpyHostToDevice),
                  "restore");
        if (nk)
            hip_check(hipMemcpy(o->d.logw + (size_t)t * o->d.cap, lw.data(), 8 * (size_t)nk, hipMemcpyHostToDevice),
                      "restore");
        auto I = download(o->d.ints + 4 * t, 4);
        I[0] = nk;
        I[1] = nk;  // slot-claim counter follows the restored count
        I[3] = ns;
        hip_check(hipMemcpy(o->d.ints + 4 * t, I.data(), 16, hipMemcpyHostToDevice), "restore");
        auto Q = download(o->d.dbl + (size_t)t * (4 + 2 * D), 4 + 2 * D);
        Q[0] = sumUprob;
        double sw = std::exp(-o->d.gamma);  // KDNorm recomputed, CommonSerialization.cpp:258-264
        for (int k = 0; k < nk; k++) sw += std::exp(lw[k]);
        Q[1] = sw;
        for (int d = 0; d < D; d++) {
            Q[4 + d] = mean[d];
            Q[4 + D + d] = M2[d];
        }
        hip_check(hipMemcpy(o->d.dbl + (size_t)t * (4 + 2 * D), Q.data(), 8 * Q.size(), hipMemcpyHostToDevice),
                  "restore");
        o->depositMirror[t] = nk;
        o->nSamplesMirror[t] = ns;
    }
    if (p >= end) return;
    count("ABMD", 0);
    if (p >= end) return;
    count("MetaD", (int)grids_.size());
    for (auto* g : grids_) {
        const int t = g->d.walkerStride ? w : 0;
        const int nd = i32();
        std::vector<double> gr(g->d.total);
        get(gr.data(), 8 * gr.size());
        g->numDeposited[t] = nd;
        hip_check(hipMemcpy(g->d.grid + (size_t)t * g->d.total, gr.data(), 8 * gr.size(), hipMemcpyHostToDevice),
                  "restore");
    }
    if (p >= end) return;
    count("PBMetaD", (int)pbs_.size());
    for (auto& ph : pbs_) {
        const int t = ph.T == 1 ? 0 : w;
        count("PBMetaD sub-grid", ph.N);
        std::vector<double> gr((size_t)ph.stride);
        size_t off = 0;
        int nd = 0;
        for (int i = 0; i < ph.N; i++) {
            nd = i32();
            get(gr.data() + off, 8 * (size_t)ph.npts[i]);
            off += ph.npts[i];
        }
        ph.numDeposited[t] = nd;
        hip_check(hipMemcpy(dPbGrid_.as<double>() + ph.base + (size_t)t * ph.stride, gr.data(), 8 * gr.size(),
                            hipMemcpyHostToDevice),
                  "restore");
    }
    if (p >= end) return;
    count("external", 0);
    if (p >= end) return;
    (void)i32();  // linear restraints: stateless, count not checked (CommonSerialization.cpp:320-326)
    if (p >= end) return;
    (void)i32();  // walls
    if (p >= end) return;
    int noe = 0, nmt = 0;
    for (auto& it : smallBias_) {
        noe += it.type == BIAS_OPES_EXPANDED;
        nmt += it.type == BIAS_OPES_MULTITHERMAL;
    }
    const size_t nsb = std::max((int)smallBias_.size(), 1);
    auto scal = download(dSmallScal_.as<double>() + (size_t)w * nsb * 4, nsb * 4);
    auto var = download(dSmallVar_.as<double>() + (size_t)w * smallVarPerWalker_, smallVarPerWalker_);
    count("OPES_EXPANDED", noe);
    for (size_t i = 0; i < smallBias_.size(); i++)
        if (smallBias_[i].type == BIAS_OPES_EXPANDED) {
            scal[4 * i] = f64();
            scal[4 * i + 1] = i32();
        }
    if (p < end) {
        count("ext-Lagrangian", 0);
        count("EDS", 0);
        count("MaxEnt", 0);
        count("multithermal", nmt);
        for (size_t i = 0; i < smallBias_.size(); i++)
            if (smallBias_[i].type == BIAS_OPES_MULTITHERMAL) {
                const auto& it = smallBias_[i];
                if (i32() != it.nState) fail("GLUED bias-state: multithermal state count mismatch");
                get(&var[it.stateOff], 8 * (size_t)it.nState);
                scal[4 * i + 2] = f64();
                scal[4 * i + 3] = f64();
            }
    }
    hip_check(hipMemcpy(dSmallScal_.as<double>() + (size_t)w * nsb * 4, scal.data(), 8 * scal.size(),
                        hipMemcpyHostToDevice),
              "restore");
    hip_check(hipMemcpy(dSmallVar_.as<double>() + (size_t)w * smallVarPerWalker_, var.data(), 8 * var.size(),
                        hipMemcpyHostToDevice),
              "restore");
    if ((size_t)(end - p) >= 8 && std::memcmp(p, "HILL", 4) == 0) {
        p += 4;
        count("hill METAD", (int)hills_.size());
        for (auto* h : hills_) {
            const int n = i32(), D = i32();
            if (D != h->d.D) fail("GLUED bias-state: hill dimension mismatch");
            if (n < 0 || n > h->d.cap) fail("GLUED bias-state: hill count out of range");
            std::vector<double> c((size_t)n * D), hh(n);
            get(c.data(), 8 * c.size());
            get(hh.data(), 8 * hh.size());
            setHills(w, h->spec, c.data(), hh.data(), n);
        }
    }
}

}  // namespace glued
