/*
 * glued.h -- C-ABI of the MI355X-native GLUED enhanced-sampling step.
 *
 * This is the drop-in boundary for the path "collective-variable evaluation ->
 * metadynamics / OPES bias evaluation and deposition -> chain-rule force scatter
 * into a 2^32 fixed-point force buffer".  Each entry point names the
 * MarvinTaterra/GluedMD interface it replaces (file:line in that repository).
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero code on failure; the
 *    message is available from glued_last_error() (thread-local).  No C++
 *    exception crosses this boundary.  (The reference throws OpenMMException
 *    from the same validation points.)
 *  - Device memory is owned by the caller for inputs and outputs: posq is
 *    float4 (x, y, z, charge) in GPU atom order (OpenMM's layout); forces are
 *    int64 field-major blocks of 3 * padded entries per walker, 2^32 fixed
 *    point, accumulated with atomics (OpenMM's long force buffer).
 *  - All device work is enqueued on the caller's stream (a hipStream_t passed
 *    as void*; NULL = default stream).  The only host synchronisations are
 *    the read-backs and the energy return of glued_execute.
 *  - No host threads are created.  A handle is used by one host thread at a
 *    time.
 */
#ifndef GLUED_H_
#define GLUED_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glued_force* glued_handle;

/* CV type codes: openmmapi/include/GluedForce.h:27-73 (same numbers). */
enum {
    GLUED_CV_DISTANCE = 1,
    GLUED_CV_ANGLE = 2,
    GLUED_CV_DIHEDRAL = 3,
    GLUED_CV_COM_DISTANCE = 4,
    GLUED_CV_GYRATION = 5,
    GLUED_CV_COORDINATION = 6, /* params [r0, n, m] + optional [cutoff] (extension) */
    GLUED_CV_RMSD = 7,         /* params 3N reference xyz + optional [fit] (1 = Kabsch, extension) */
    GLUED_CV_PATH = 10,
    GLUED_CV_POSITION = 11,
    GLUED_CV_DRMSD = 12,
    GLUED_CV_CONTACTMAP = 13,
    GLUED_CV_ENERGY = 23 /* value and Jacobian supplied by glued_set_energy_input */
};

/* Bias type codes: openmmapi/include/GluedForce.h:76-95 (same numbers), plus
 * GLUED_BIAS_METAD_HILLS, a hill-list well-tempered metadynamics with no
 * counterpart in the reference (params [h0, sigma*D, gamma, kT, period*D],
 * int params [pace, maxHills]; period 0 = non-periodic). */
enum {
    GLUED_BIAS_HARMONIC = 1,     /* params [k, s0] per CV; int params [periodic] (optional) */
    GLUED_BIAS_METAD = 3,
    GLUED_BIAS_PBMETAD = 4,      /* params [h0, gamma, kT, (sigma, min, max) per CV];
                                    int params [pace, (bins, periodic) per CV] */
    GLUED_BIAS_OPES = 5,
    GLUED_BIAS_UPPER_WALL = 8,   /* params [at, kappa, eps, n] per CV */
    GLUED_BIAS_LOWER_WALL = 9,   /* params [at, kappa, eps, n] per CV */
    GLUED_BIAS_LINEAR = 10,      /* params [k] per CV */
    GLUED_BIAS_OPES_EXPANDED = 11,
    GLUED_BIAS_OPES_MULTITHERMAL = 15,
    GLUED_BIAS_METAD_HILLS = 101
};

/* Thread-local message of the last failure. */
const char* glued_last_error(void);

/* ---- lifecycle and specification -------------------------------------- */

/* Replaces GluedForce::GluedForce (openmmapi/src/GluedForce.cpp:12-14). */
int glued_create(glued_handle* out);
/* Replaces the Context-owned kernel's destruction (GluedForceImpl.cpp:31-34). */
int glued_destroy(glued_handle h);
/* Replaces GluedForce::setUsesPeriodicBoundaryConditions (GluedForce.h:163). */
int glued_set_pbc(glued_handle h, int yes);
/* Replaces GluedForce::setTemperature (GluedForce.cpp:121-124). */
int glued_set_temperature(glued_handle h, double kelvin);
/* Replaces GluedForce::addCollectiveVariable (GluedForce.cpp:44-62): returns
 * the first value index (PATH occupies two). */
int glued_add_cv(glued_handle h, int type, const int* atoms, int nAtoms, const double* params, int nParams,
                 int* firstValueIdx);
/* Replaces GluedForce::addBias (GluedForce.cpp:82-107). */
int glued_add_bias(glued_handle h, int type, const int* cvIdx, int nCv, const double* params, int nParams,
                   const int* iparams, int nIparams, int* biasIdx);
/* Replaces GluedForce::getNumCollectiveVariables (GluedForce.cpp:64-66). */
int glued_num_cv_values(glued_handle h, int* n);
/* Replaces GluedForce::setTestBiasGradients (GluedForce.cpp:186-189): inject
 * dU/dCV, skipping bias evaluation (CommonGluedKernels.cpp:4670-4671). */
int glued_set_test_bias_gradients(glued_handle h, const double* grads, int n);
/* One bias state shared by every walker of the handle (the within-context
 * form of the reference's multi-walker sharing, CommonGluedKernels.cpp
 * :2912-2933).  Call before initialization. */
int glued_share_bias(glued_handle h, int biasIdx, int shared);
/* Replaces GluedForce::getMultiWalkerPtrs (GluedForce.cpp:257-262,
 * GluedKernels.h:89-96): the device addresses of bias biasIdx's table, for
 * glued_redirect_to_primary_bias on another handle of the same device.
 * METAD: {grid}; METAD_HILLS: {centres, heights, count}; OPES: {centres,
 * widths, log-weights, counts (numKernels, numAllocated, steps, samples),
 * sums (sum_uprob, KDNorm, sum_w, sum_w2, Welford mean and M2)}.  *count
 * receives the number of addresses (at most capacity are written).  The bias
 * must have one table (a one-walker handle, or glued_share_bias).  After
 * initialization. */
int glued_get_multiwalker_ptrs(glued_handle h, int biasIdx, int64_t* ptrs, int capacity, int* count);
/* Replaces GluedForce::setMultiWalkerPtrs / redirectToPrimaryBias
 * (GluedForce.cpp:264-270, CommonGluedKernels.cpp:5234-5262): this handle's
 * table of bias biasIdx becomes the table at ptrs (from
 * glued_get_multiwalker_ptrs of a handle with the same bias definition).
 * The whole table is shared, its OPES sums included, so handles stepped in
 * turn on one stream give the results of one handle whose walkers share the
 * table.  Each address is checked to lie in a device allocation large enough
 * for this bias's table.  The primary handle must outlive this one; handles
 * sharing a table must not run deposits concurrently (one stream, or ordered
 * streams). */
int glued_redirect_to_primary_bias(glued_handle h, int biasIdx, const int64_t* ptrs, int count);

/* ---- initialization ---------------------------------------------------- */

/* Replaces CommonCalcGluedForceKernel::initialize (CommonGluedKernels.cpp:
 * 4608-4644) with its validation (:3879-3927).  One walker; masses may be
 * NULL (all 1). */
int glued_initialize(glued_handle h, int numParticles, int paddedNumParticles, const double* masses, int nMasses);
/* Batched form (no reference counterpart; the reference runs one Context per
 * walker): W walkers of uneven size, laid out back to back in posq / forces
 * (walker w starts at sum_{v<w} padded_v atoms). */
int glued_initialize_batch(glued_handle h, int numWalkers, const int* numParticles, const int* paddedNumParticles,
                           const double* masses, int nMasses);
/* Replaces rebuildGpuAtomIndices (CommonCVKernels.cpp:113-317). */
int glued_set_atom_order(glued_handle h, int walker, const int* gpuToUser, int n);
/* Replaces setPeriodicBoxArgs (CommonGluedKernels.cpp:4676-4734): box vectors
 * a, b, c (9 doubles, OpenMM reduced form). */
int glued_set_box(glued_handle h, int walker, const double* box9);
/* CV_ENERGY input: U and the walker's external force (device FP32,
 * field-major 3*padded); the energy CV's Jacobian is -F (CommonGluedKernels.cpp:
 * 2471-2503 copies -F the same way from an inner Context). */
int glued_set_energy_input(glued_handle h, int walker, double U, const float* forceField, void* stream);

/* ---- the step ------------------------------------------------------------ */

/* Replaces CalcGluedForceKernel::execute (GluedKernels.h:48-50,
 * CommonGluedKernels.cpp:4646-5048) for one walker; posq and forces point at
 * that walker's block.  *energy receives the summed bias energy when
 * includeEnergy (one host synchronisation), else 0. */
int glued_execute(glued_handle h, void* stream, int walker, const void* posq, int64_t* forces, int includeForces,
                  int includeEnergy, double* energy);
/* glued_execute (without the energy read-back) for walkers 0..W-1 in turn,
 * looped inside the library: posq and forces are the batch bases.  The
 * reference's structure for several contexts (one CalcGluedForceKernel::
 * execute per context, CommonGluedKernels.cpp:4646-5048) with no per-walker
 * foreign-call overhead; results equal W glued_execute calls. */
int glued_execute_plain_all(glued_handle h, void* stream, const void* posq, int64_t* forces, int includeForces);
/* Batched form: every walker in one launch per stage; useGraph = 1 replays
 * the step as one captured hipGraph. */
int glued_execute_batch(glued_handle h, void* stream, const void* posq, int64_t* forces, int includeForces,
                        int useGraph);
/* One stage of the batched step for every walker (0 = CVs, 1 = biases,
 * 2 = dU/dCV + force scatter, 3 = only the streamed hill / OPES-kernel sums,
 * 4 = only the force scatter: the combined dense energy-CV + atom-slot pass
 * where the batched step uses it, else the dense energy-CV scatter), for
 * per-stage timing; no reference
 * counterpart. */
int glued_execute_stage(glued_handle h, void* stream, const void* posq, int64_t* forces, int stage);
/* Replaces CalcGluedForceKernel::updateState (GluedKernels.h:52,
 * CommonGluedKernels.cpp:5050-5157) with GluedForceImpl's once-per-step guard
 * (GluedForceImpl.cpp:100-109); all walkers. */
int glued_update_state(glued_handle h, void* stream, int step);
/* One MD step for every walker: glued_update_state(step) followed by
 * glued_execute_batch (the order OpenMM uses: updateContextState, then the
 * force evaluation; GluedForceImpl.cpp:100-109), replayed as one hipGraph per
 * deposit pattern.  Results and host-side state are those of the two calls;
 * no reference counterpart (the reference launches each kernel separately). */
int glued_step_batch(glued_handle h, void* stream, int step, const void* posq, int64_t* forces, int includeForces);

/* ---- read-backs ---------------------------------------------------------- */

/* Replaces GluedForce::getLastCVValues (GluedForce.cpp:206-213). */
int glued_get_cv_values(glued_handle h, int walker, double* out, int n);
/* Per-bias energies of the last evaluation; their sum replaces
 * GluedForce::getLastBias (GluedForce.cpp:215-222). */
int glued_get_bias_energies(glued_handle h, int walker, double* out, int n);
/* dU/dCV of the last evaluation (the reference's cvBiasGradients). */
int glued_get_dudcv(glued_handle h, int walker, double* out, int n);
/* Pairs inside the cutoff of the coordIdx-th coordination CV (extension). */
int glued_get_coordination_pairs(glued_handle h, int walker, int coordIdx, long long* npairs);
/* Replaces GluedForce::getBiasStateBytes / setBiasStateBytes
 * (GluedForce.cpp:133-143, CommonSerialization.cpp:20-410), 'GPUS' v1. */
int glued_get_bias_state(glued_handle h, int walker, char* buf, size_t cap, size_t* size);
int glued_set_bias_state(glued_handle h, int walker, const char* buf, size_t size);
/* Replaces GluedForce::getOPESMetrics (CommonGluedKernels.cpp:5164-5210):
 * out = {zed, rct, nker, neff}. */
int glued_get_opes_metrics(glued_handle h, int walker, int opesIdx, double* out4);
/* Replaces GluedForce::getKernelSigmas (CommonGluedKernels.cpp:5212-5232). */
int glued_get_kernel_sigmas(glued_handle h, int walker, int opesIdx, float* out, int cap, int* n);
/* Hill list of a GLUED_BIAS_METAD_HILLS bias (extension): centres kernel-major
 * (n x D), heights n. */
int glued_set_hills(glued_handle h, int walker, int biasIdx, const double* centers, const double* heights, int n);
int glued_get_hills(glued_handle h, int walker, int biasIdx, double* centers, double* heights, int cap, int* n);
/* Whole MetaD grid (the grid section of the state blob). */
int glued_set_grid(glued_handle h, int walker, int biasIdx, const double* grid, int n);
/* ---- autograd support (no reference counterpart) ------------------------ */

/* Device-to-device copy of per-walker state: 0 = CV values [W][V],
 * 1 = per-bias gradients dV_b/ds_d [W][G], 2 = bias energies [W][B],
 * 3 = dU/dCV [W][V], 4 = (ds/dlambda, dz/dlambda) of every path CV
 * [W][nPath][2]. */
int glued_copy_buffer(glued_handle h, int which, void* dst, size_t bytes, void* stream);
/* Hill heights of a GLUED_BIAS_METAD_HILLS bias from device memory. */
int glued_set_hill_heights(glued_handle h, int walker, int biasIdx, const double* dHeights, int n, void* stream);
/* lambda of the pathIdx-th path CV (FP64). */
int glued_set_path_lambda(glued_handle h, int pathIdx, double lambda);
/* Widths of a GLUED_BIAS_METAD_HILLS bias. */
int glued_set_hill_sigma(glued_handle h, int biasIdx, const double* sigma, int D);
/* Adjoint of a hill bias: given g_w = dL/dV_w (device, [W]) writes
 * dL/dh_k (device, [T][cap], T = 1 shared or W private tables) and
 * dL/dsigma_d (device, [D]) for the CV values of the last evaluation. */
int glued_hill_adjoint(glued_handle h, void* stream, int biasIdx, const double* dGradV, double* dGradHeights,
                       double* dGradSigma);
/* glued_hill_adjoint without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc): the
 * per-block partials live in a fixed 512 KiB device array of the library
 * (65,536 doubles per GPU, held until the process ends), so this call makes
 * no hipMalloc / hipFree.  A bias needing more, ceil(cap / 256) * T * D
 * doubles (T = 1 shared or W private tables), is rejected with an error.
 * Same arguments and results as glued_hill_adjoint, bit for bit.  Calls
 * that overlap on one GPU share the array and race. */
int glued_hill_adjoint_noalloc(glued_handle h, void* stream, int biasIdx, const double* dGradV,
                               double* dGradHeights, double* dGradSigma);
/* Adjoint of an OPES bias with respect to its kernel table (no reference
 * counterpart; the function differentiated is opesEvalBias,
 * CommonGluedKernels.cpp:2642-2723): given g_w = dL/dV_w (device, [W]) writes
 * dL/d(centres) and dL/d(sigmas) (device, [T][cap][D] kernel-major),
 * dL/d(log-weights) ([T][cap]) and dL/d(sum_uprob) ([T]), T = 1 shared or W
 * private tables, for the CV values and energies of the last evaluation;
 * sum_uprob is treated as the stored input it is in the eval. */
int glued_opes_adjoint(glued_handle h, void* stream, int biasIdx, const double* dGradV, double* dGradCenters,
                       double* dGradSigmas, double* dGradLogW, double* dGradSumUprob);
/* The kernel table of an OPES bias set directly (host, kernel-major centres
 * and sigmas [n][D], log-weights [n]) with its sum_uprob; kernel count and
 * slot-claim counter follow n.  No reference counterpart (the reference loads
 * tables only from the state blob, CommonSerialization.cpp:180-410). */
int glued_set_opes_kernels(glued_handle h, int walker, int biasIdx, const double* centers, const double* sigmas,
                           const double* logw, int n, double sumUprob);
/* Reference frames of the pathIdx-th path CV (host, [N][M][3], stored as
 * FP32 like the CV's parameters); n = N*M*3.  No reference counterpart: the
 * reference fixes the frames at creation (CommonGluedKernels.cpp:4011-4024). */
int glued_set_path_frames(glued_handle h, int pathIdx, const double* frames, int n);
/* Adjoint of the pathIdx-th path CV with respect to its frames: given
 * g_w = dL/dV_w (device, [W]) writes dL/dr (device, [N][M][3]) for the
 * positions, CV values and dU/dCV of the last evaluation.  No reference
 * counterpart (the reference has no frame derivatives). */
int glued_path_frames_adjoint(glued_handle h, void* stream, int pathIdx, const double* dGradV, double* dGradFrames);

/* Launch shapes chosen by the autotuner (no reference counterpart): the
 * path MSD block target, the path row phases and the coordination cell
 * workgroup size, how they were chosen (measured at the first batched
 * evaluation, cached for the plan shape, or default) and the measured CV
 * stage times; a NUL-terminated text of *len characters (buf may be NULL to
 * ask for the length). */
int glued_get_tuning(glued_handle h, char* buf, size_t cap, size_t* len);
/* Plan sizes: Jacobian rows and distinct atoms of one walker. */
int glued_plan_info(glued_handle h, int* rows, int* atoms);

#ifdef __cplusplus
}
#endif

#endif /* GLUED_H_ */
