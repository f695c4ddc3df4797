// C-ABI of the GLUED step (declared in include/glued.h).  Converts every C++
// exception into a status code and a thread-local message.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include "../../include/glued.h"
#include "engine.hpp"

using glued::Engine;

struct glued_force {
    Engine e;
};

static thread_local std::string g_err;

// ---- workspace of glued_hill_adjoint_noalloc -------------------------------
// The per-block sigma partials of the hill adjoint live in this file-scope
// device array instead of engine-allocated memory, so that the entry point
// never allocates.  Every GPU the library runs on holds its own copy (512 KiB)
// until the process ends; its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) at the handle's
// initialisation, or at the entry point's first call on a device.  In host
// code the name `scratch` is only a host-side stand-in: kernels and copies get
// the looked-up address, never the name.  Every call on one device shares the
// array: calls that overlap on one GPU race.  Nothing in it needs clearing:
// each call's first kernel writes every partial its second kernel reads.
constexpr size_t kScratchDoubles = 65536;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static double* scratch_for_current_device()
{
    int dev = -1;
    glued::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw glued::Error("hill_adjoint_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw glued::Error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return static_cast<double*>(g_scratch_ptr[dev]);
}

template <class F>
static int guard(F&& f)
{
    try {
        f();
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    } catch (...) {
        g_err = "unknown error";
        return 1;
    }
}

static Engine& E(glued_handle h)
{
    if (!h) throw glued::Error("null handle");
    return h->e;
}

template <class T>
static void copy_out(const std::vector<T>& v, T* out, int n, const char* what)
{
    if ((int)v.size() > n) throw glued::Error(std::string(what) + ": output buffer too small");
    if (!v.empty()) std::memcpy(out, v.data(), sizeof(T) * v.size());
}

#pragma GCC visibility push(default)
extern "C" {

const char* glued_last_error(void) { return g_err.c_str(); }

int glued_create(glued_handle* out)
{
    return guard([&] {
        if (!out) throw glued::Error("null output pointer");
        *out = new glued_force();
    });
}

int glued_destroy(glued_handle h)
{
    return guard([&] { delete h; });
}

int glued_set_pbc(glued_handle h, int yes)
{
    return guard([&] { E(h).pbc = yes != 0; });
}

int glued_set_temperature(glued_handle h, double kelvin)
{
    return guard([&] { E(h).temperature = kelvin; });
}

int glued_add_cv(glued_handle h, int type, const int* atoms, int nAtoms, const double* params, int nParams,
                 int* firstValueIdx)
{
    return guard([&] {
        int v = E(h).addCV(type, atoms, nAtoms, params, nParams);
        if (firstValueIdx) *firstValueIdx = v;
    });
}

int glued_add_bias(glued_handle h, int type, const int* cvIdx, int nCv, const double* params, int nParams,
                   const int* iparams, int nIparams, int* biasIdx)
{
    return guard([&] {
        int b = E(h).addBias(type, cvIdx, nCv, params, nParams, iparams, nIparams);
        if (biasIdx) *biasIdx = b;
    });
}

int glued_num_cv_values(glued_handle h, int* n)
{
    return guard([&] { *n = E(h).numValues(); });
}

int glued_set_test_bias_gradients(glued_handle h, const double* grads, int n)
{
    return guard([&] { E(h).setTestBiasGradients(grads, n); });
}

int glued_share_bias(glued_handle h, int biasIdx, int shared)
{
    return guard([&] { E(h).setShared(biasIdx, shared); });
}

int glued_get_multiwalker_ptrs(glued_handle h, int biasIdx, int64_t* ptrs, int capacity, int* count)
{
    return guard([&] {
        const std::vector<long long> p = E(h).multiWalkerPtrs(biasIdx);
        if (count) *count = (int)p.size();
        for (int k = 0; k < (int)p.size() && k < capacity; k++) ptrs[k] = p[k];
    });
}

int glued_redirect_to_primary_bias(glued_handle h, int biasIdx, const int64_t* ptrs, int count)
{
    return guard([&] {
        if (count < 0 || (count > 0 && !ptrs)) throw std::invalid_argument("redirect_bias: no pointers");
        std::vector<long long> p(ptrs, ptrs + count);
        E(h).redirectBias(biasIdx, p.data(), count);
    });
}

int glued_initialize(glued_handle h, int numParticles, int paddedNumParticles, const double* masses, int nMasses)
{
    return guard([&] {
        E(h).initialize(1, &numParticles, &paddedNumParticles, nMasses, masses);
        (void)scratch_for_current_device();
    });
}

int glued_initialize_batch(glued_handle h, int numWalkers, const int* numParticles, const int* paddedNumParticles,
                           const double* masses, int nMasses)
{
    return guard([&] {
        E(h).initialize(numWalkers, numParticles, paddedNumParticles, nMasses, masses);
        (void)scratch_for_current_device();
    });
}

int glued_set_atom_order(glued_handle h, int walker, const int* gpuToUser, int n)
{
    return guard([&] { E(h).setAtomOrder(walker, gpuToUser, n); });
}

int glued_set_box(glued_handle h, int walker, const double* box9)
{
    return guard([&] { E(h).setBox(walker, box9); });
}

int glued_set_energy_input(glued_handle h, int walker, double U, const float* forceField, void* stream)
{
    return guard([&] { E(h).setEnergyInput(walker, U, forceField, (hipStream_t)stream); });
}

int glued_execute(glued_handle h, void* stream, int walker, const void* posq, int64_t* forces, int includeForces,
                  int includeEnergy, double* energy)
{
    return guard([&] {
        double e = E(h).execute((hipStream_t)stream, walker, (const float4*)posq, (long long*)forces,
                                includeForces != 0, includeEnergy != 0);
        if (energy) *energy = e;
    });
}

int glued_execute_plain_all(glued_handle h, void* stream, const void* posq, int64_t* forces, int includeForces)
{
    return guard([&] { E(h).executePlainAll((hipStream_t)stream, (const float4*)posq, (long long*)forces,
                                            includeForces != 0); });
}

int glued_execute_batch(glued_handle h, void* stream, const void* posq, int64_t* forces, int includeForces,
                        int useGraph)
{
    return guard([&] {
        E(h).executeBatch((hipStream_t)stream, (const float4*)posq, (long long*)forces, includeForces != 0,
                          useGraph != 0);
    });
}

int glued_execute_stage(glued_handle h, void* stream, const void* posq, int64_t* forces, int stage)
{
    return guard([&] { E(h).executeStage((hipStream_t)stream, (const float4*)posq, (long long*)forces, stage); });
}

int glued_update_state(glued_handle h, void* stream, int step)
{
    return guard([&] { E(h).updateState((hipStream_t)stream, step); });
}

int glued_step_batch(glued_handle h, void* stream, int step, const void* posq, int64_t* forces, int includeForces)
{
    return guard([&] {
        E(h).stepBatch((hipStream_t)stream, step, (const float4*)posq, (long long*)forces, includeForces != 0);
    });
}

int glued_get_cv_values(glued_handle h, int walker, double* out, int n)
{
    return guard([&] { copy_out(E(h).cvValues(walker), out, n, "get_cv_values"); });
}

int glued_get_bias_energies(glued_handle h, int walker, double* out, int n)
{
    return guard([&] { copy_out(E(h).biasEnergies(walker), out, n, "get_bias_energies"); });
}

int glued_get_dudcv(glued_handle h, int walker, double* out, int n)
{
    return guard([&] { copy_out(E(h).dUdCV(walker), out, n, "get_dudcv"); });
}

int glued_get_coordination_pairs(glued_handle h, int walker, int coordIdx, long long* npairs)
{
    return guard([&] { *npairs = E(h).coordinationPairs(walker, coordIdx); });
}

int glued_get_bias_state(glued_handle h, int walker, char* buf, size_t cap, size_t* size)
{
    return guard([&] {
        auto b = E(h).getState(walker);
        if (size) *size = b.size();
        if (buf) {
            if (cap < b.size()) throw glued::Error("get_bias_state: buffer too small");
            std::memcpy(buf, b.data(), b.size());
        }
    });
}

int glued_set_bias_state(glued_handle h, int walker, const char* buf, size_t size)
{
    return guard([&] { E(h).setState(walker, std::vector<char>(buf, buf + size)); });
}

int glued_get_opes_metrics(glued_handle h, int walker, int opesIdx, double* out4)
{
    return guard([&] {
        auto m = E(h).opesMetrics(walker, opesIdx);
        for (int i = 0; i < 4; i++) out4[i] = m[i];
    });
}

int glued_get_kernel_sigmas(glued_handle h, int walker, int opesIdx, float* out, int cap, int* n)
{
    return guard([&] {
        auto s = E(h).kernelSigmas(walker, opesIdx);
        if (n) *n = (int)s.size();
        if (out) copy_out(s, out, cap, "get_kernel_sigmas");
    });
}

int glued_set_hills(glued_handle h, int walker, int biasIdx, const double* centers, const double* heights, int n)
{
    return guard([&] { E(h).setHills(walker, biasIdx, centers, heights, n); });
}

int glued_get_hills(glued_handle h, int walker, int biasIdx, double* centers, double* heights, int cap, int* n)
{
    return guard([&] {
        int k = E(h).getHills(walker, biasIdx, centers, heights, cap);
        if (n) *n = k;
    });
}

int glued_set_grid(glued_handle h, int walker, int biasIdx, const double* grid, int n)
{
    return guard([&] { E(h).setGrid(walker, biasIdx, grid, n); });
}

int glued_copy_buffer(glued_handle h, int which, void* dst, size_t bytes, void* stream)
{
    return guard([&] { E(h).copyBuffer(which, dst, bytes, (hipStream_t)stream); });
}

int glued_set_hill_heights(glued_handle h, int walker, int biasIdx, const double* dHeights, int n, void* stream)
{
    return guard([&] { E(h).setHillHeightsDevice(walker, biasIdx, dHeights, n, (hipStream_t)stream); });
}

int glued_set_path_lambda(glued_handle h, int pathIdx, double lambda)
{
    return guard([&] { E(h).setPathLambda(pathIdx, lambda); });
}

int glued_set_hill_sigma(glued_handle h, int biasIdx, const double* sigma, int D)
{
    return guard([&] { E(h).setHillSigma(biasIdx, sigma, D); });
}

int glued_hill_adjoint(glued_handle h, void* stream, int biasIdx, const double* dGradV, double* dGradHeights,
                       double* dGradSigma)
{
    return guard([&] { E(h).hillAdjoint((hipStream_t)stream, biasIdx, dGradV, dGradHeights, dGradSigma); });
}

int glued_hill_adjoint_noalloc(glued_handle h, void* stream, int biasIdx, const double* dGradV,
                               double* dGradHeights, double* dGradSigma)
{
    return guard([&] {
        Engine& e = E(h);
        const size_t need = e.hillAdjointPartials(biasIdx);
        if (need > kScratchDoubles)
            throw glued::Error("hill_adjoint_noalloc: this bias needs " + std::to_string(need) +
                               " doubles of workspace, more than the " + std::to_string(kScratchDoubles) +
                               " of the library's fixed array; use glued_hill_adjoint");
        e.hillAdjoint((hipStream_t)stream, biasIdx, dGradV, dGradHeights, dGradSigma, scratch_for_current_device(),
                      kScratchDoubles);
    });
}

int glued_opes_adjoint(glued_handle h, void* stream, int biasIdx, const double* dGradV, double* dGradCenters,
                       double* dGradSigmas, double* dGradLogW, double* dGradSumUprob)
{
    return guard([&] {
        E(h).opesAdjoint((hipStream_t)stream, biasIdx, dGradV, dGradCenters, dGradSigmas, dGradLogW, dGradSumUprob);
    });
}

int glued_set_opes_kernels(glued_handle h, int walker, int biasIdx, const double* centers, const double* sigmas,
                           const double* logw, int n, double sumUprob)
{
    return guard([&] { E(h).setOpesKernels(walker, biasIdx, centers, sigmas, logw, n, sumUprob); });
}

int glued_set_path_frames(glued_handle h, int pathIdx, const double* frames, int n)
{
    return guard([&] { E(h).setPathFrames(pathIdx, frames, n); });
}

int glued_path_frames_adjoint(glued_handle h, void* stream, int pathIdx, const double* dGradV, double* dGradFrames)
{
    return guard([&] { E(h).pathFramesAdjoint((hipStream_t)stream, pathIdx, dGradV, dGradFrames); });
}

int glued_get_tuning(glued_handle h, char* buf, size_t cap, size_t* len)
{
    return guard([&] {
        const std::string r = E(h).tuningReport();
        if (len) *len = r.size();
        if (buf && cap) {
            const size_t n = std::min(cap - 1, r.size());
            std::memcpy(buf, r.data(), n);
            buf[n] = 0;
        }
    });
}

int glued_plan_info(glued_handle h, int* rows, int* atoms)
{
    return guard([&] {
        if (rows) *rows = E(h).jacobianRows();
        if (atoms) *atoms = E(h).atomSlots();
    });
}

}  // extern "C"
#pragma GCC visibility pop
