// Bias evaluation and deposition kernels for gfx950.
//
// Chunked sums (hill-list metadynamics, OPES kernels) stream the bias state
// once per evaluation: each workgroup reduces one chunk to FP64 partials;
// the partials are then added in a fixed order (wave-parallel, see
// reduce_partials) -- deterministic -- and V and dV/ds written.  Per-hill
// exponents are formed in FP64 with unfused operations so the E <= 8
// truncation predicate is bit-identical to the FP64 oracle; exp() runs in FP32
// (its relative error bounds each term by 0.37 * 2^-24 of its height).
#include "glued_internal.h"
#include <hip/hip_runtime.h>
#include "scatter_dev.h"
#include "wave_reduce.h"
#include "cv_small_dev.h"
#include <cmath>

namespace glued {

__device__ inline double wsum(double v) { return wave_sum1(v); }

// ---------------------------------------------------------------------------
// Grid MetaD evaluation: CommonGluedKernels.cpp:3574-3636.
// ---------------------------------------------------------------------------
template <int D>
__device__ void grid_eval_t(const GridBiasDev& g, const double* grid, const double* s, double* V, double* dV)
{
    int lo[D], hi[D];
    double al[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        double fr = (s[d] - g.origin[d]) * g.invSpacing[d];
        const int N = g.npts[d];
        if (N <= 1) {
            lo[d] = hi[d] = 0;
            al[d] = 0.0;
        } else if (g.periodic[d]) {
            fr = fr - floor(fr / (double)N) * (double)N;
            if (fr < 0.0) fr += N;
            lo[d] = (int)fr;
            if (lo[d] >= N) lo[d] = N - 1;
            al[d] = fr - lo[d];
            hi[d] = (lo[d] + 1) % N;
        } else {
            if (fr < 0.0) fr = 0.0;
            if (fr > (double)(N - 1)) fr = (double)(N - 1);
            lo[d] = (int)fr;
            if (lo[d] >= N - 1) lo[d] = N - 2;
            al[d] = fr - lo[d];
            hi[d] = lo[d] + 1;
        }
    }
    double v = 0.0, gr[D];
#pragma unroll
    for (int d = 0; d < D; d++) gr[d] = 0.0;
#pragma unroll
    for (int c = 0; c < (1 << D); c++) {
        int idx = 0;
#pragma unroll
        for (int d = 0; d < D; d++) idx += (((c >> d) & 1) ? hi[d] : lo[d]) * g.strides[d];
        const double val = grid[idx];
        double w = 1.0;
#pragma unroll
        for (int d = 0; d < D; d++) w *= ((c >> d) & 1) ? al[d] : (1.0 - al[d]);
        v += w * val;
#pragma unroll
        for (int d = 0; d < D; d++) {
            double wo = 1.0;
#pragma unroll
            for (int e = 0; e < D; e++)
                if (e != d) wo *= ((c >> e) & 1) ? al[e] : (1.0 - al[e]);
            gr[d] += (((c >> d) & 1) ? 1.0 : -1.0) * wo * val;
        }
    }
    *V = v;
#pragma unroll
    for (int d = 0; d < D; d++) dV[d] = gr[d] * g.invSpacing[d];
}

// Grid MetaD evaluation: CommonGluedKernels.cpp:3574-3636 (D <= 3).
__device__ inline void grid_eval(const GridBiasDev& g, const double* grid, const double* cv, double* be, double* bg)
{
    double s[3], dV[3], V;
    if (g.D == 1) {
        s[0] = cv[g.cv[0]];
        grid_eval_t<1>(g, grid, s, &V, dV);
        bg[0] = dV[0];
    } else if (g.D == 2) {
        s[0] = cv[g.cv[0]];
        s[1] = cv[g.cv[1]];
        grid_eval_t<2>(g, grid, s, &V, dV);
        bg[0] = dV[0];
        bg[1] = dV[1];
    } else {
        s[0] = cv[g.cv[0]];
        s[1] = cv[g.cv[1]];
        s[2] = cv[g.cv[2]];
        grid_eval_t<3>(g, grid, s, &V, dV);
        bg[0] = dV[0];
        bg[1] = dV[1];
        bg[2] = dV[2];
    }
    *be = V;
}

// ---------------------------------------------------------------------------
// One-thread biases: grid MetaD eval, OPES expanded (:3050-3090),
// multithermal (:3145-3172).  One wave per (item, walker): every lane calls
// with the same item.  Grid MetaD runs on lane 0; OPES expanded and
// multithermal give lane l the states l, l + 64, ...: the maximum is exact
// and the sums are wave butterflies in a fixed order, so the plain and the
// batched forms (both call this) agree bit for bit.
// ---------------------------------------------------------------------------
__device__ inline double wave_max1(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}

// bgL (optional, LDS): a copy of the walker's bias-gradient slots.
// Static restraints, CV d of the bias at value s: adds to V, returns dV/ds.
// Harmonic CommonGluedKernels.cpp:2506-2540 (periodic flag: the difference
// folded to the nearest 2 pi image), linear :2542-2566, walls :2568-2616
// (delta = s - at, or at - s for a lower wall; only delta > 0 counts;
// n = 1, 2 in closed form, pow() otherwise).  Unfused, as the oracle.
__device__ inline double restraint_term(const SmallBiasItem& it, const double* p, int d, double s, double& V)
{
#pragma clang fp contract(off)
    const int flag = it.bias;
    if (it.type == BIAS_HARMONIC) {
        const double TWO_PI = 6.283185307179586;
        const double k = p[2 * d];
        double ds = s - p[2 * d + 1];
        if (flag) ds -= TWO_PI * round(ds / TWO_PI);
        V += 0.5 * k * ds * ds;
        return k * ds;
    }
    if (it.type == BIAS_LINEAR) {
        V += p[d] * s;
        return p[d];
    }
    const double at = p[4 * d], kappa = p[4 * d + 1], eps = p[4 * d + 2], n = p[4 * d + 3];
    const double delta = flag ? at - s : s - at;
    if (delta <= 0.0) return 0.0;
    double powered, dPow;
    if (n == 1.0) {
        powered = delta;
        dPow = 1.0;
    } else if (n == 2.0) {
        powered = delta * delta;
        dPow = 2.0 * delta;
    } else {
        powered = pow(delta, n);
        dPow = n * pow(delta, n - 1.0);
    }
    const double e = exp(eps * delta);
    V += kappa * powered * e;
    return (flag ? -1.0 : 1.0) * kappa * (dPow + eps * powered) * e;
}

// PBMetaD, CommonGluedKernels.cpp:3697-3763 (one 1-D grid per CV, combined
// by log-sum-exp): sub-grid i interpolated linearly at s_i (periodic wrap, or
// clamped with lo <= N - 2), V_i and dV_i/ds_i; V = -kT (log sum_i
// e^((Vmin - V_i)/kT) - log N) + Vmin, dV/ds_i = w_i dV_i/ds_i with the
// softmax weights w_i, which are kept with V_i for the deposit.  Constants
// per sub-grid: {origin, 1/spacing, points, periodic, sigma, offset}.
constexpr int PB_CONST = 6;

__device__ inline void pbmetad_eval(const StepView& v, const SmallBiasState& st, const SmallBiasItem& it, int w,
                                    double* bgL)
{
#pragma clang fp contract(off)
    const double* cv = v.cvVal + (size_t)w * v.V;
    const double* c = st.stateConst + it.stateOff;
    double* var = st.stateVar + (size_t)w * st.stateVarStride + it.stateOff;  // V_i, then w_i
    const int tb = it.p[3] != 0.0 ? 0 : w;
    const double* grid = st.pbGrid + it.bias + (size_t)tb * it.nState;
    const int N = it.D;
    const double kT = it.p[1];
    double Vi[MAX_BIAS_D], gi[MAX_BIAS_D];
    for (int i = 0; i < N; i++) {
        const double* ci = c + PB_CONST * i;
        const int np = (int)ci[2];
        double frac = (cv[it.cv[i]] - ci[0]) * ci[1];
        int lo, hi;
        if (ci[3] != 0.0) {
            frac = frac - floor(frac / (double)np) * (double)np;
            if (frac < 0.0) frac += (double)np;
            lo = (int)frac;
            if (lo >= np) lo = np - 1;
            hi = (lo + 1) % np;
        } else {
            if (frac < 0.0) frac = 0.0;
            if (frac > (double)(np - 1)) frac = (double)(np - 1);
            lo = (int)frac;
            if (lo >= np - 1) lo = np - 2;
            hi = lo + 1;
        }
        const double alpha = frac - lo;
        const double* g = grid + (int)ci[5];
        const double vlo = g[lo], vhi = g[hi];
        Vi[i] = vlo * (1.0 - alpha) + vhi * alpha;
        gi[i] = (vhi - vlo) * ci[1];
    }
    double bmin = Vi[0];
    for (int i = 1; i < N; i++)
        if (Vi[i] < bmin) bmin = Vi[i];
    double ene = 0.0;
    for (int i = 0; i < N; i++) ene += exp((bmin - Vi[i]) / kT);
    v.benergy[(size_t)w * v.B + it.eslot] = -kT * (log(ene) - log((double)N)) + bmin;
    double* bg = v.bgrad + (size_t)w * v.G + it.gslot;
    for (int i = 0; i < N; i++) {
        const double wi = exp((bmin - Vi[i]) / kT) / ene;
        bg[i] = wi * gi[i];
        if (bgL) bgL[it.gslot + i] = wi * gi[i];
        var[i] = Vi[i];
        var[N + i] = wi;
    }
}

// Every CV of a restraint item: energy slot, gradient slots (and the LDS
// mirror).
__device__ inline void restraints_eval(const StepView& v, const SmallBiasState& st, const SmallBiasItem& it, int w,
                                       double* bgL)
{
    const double* cv = v.cvVal + (size_t)w * v.V;
    const double* p = st.stateConst + it.stateOff;  // parameters, FP32-rounded
    double* bg = v.bgrad + (size_t)w * v.G + it.gslot;
    double V = 0.0;
    for (int d = 0; d < it.D; d++) {
        const double g = restraint_term(it, p, d, cv[it.cv[d]], V);
        bg[d] = g;
        if (bgL) bgL[it.gslot + d] = g;
    }
    v.benergy[(size_t)w * v.B + it.eslot] = V;
}

// EXTRA: the plan has restraint or PBMetaD items (kernels of plans without
// them are compiled without that code, which would otherwise cost the
// finalize time).
template <bool EXTRA = true>
__device__ void small_bias_eval(const StepView& v, const SmallBiasState& st, const SmallBiasItem& it, int i, int w,
                                double* bgL = nullptr)
{
    const int lane = threadIdx.x & 63;
    const double* cv = v.cvVal + (size_t)w * v.V;
    double* bg = v.bgrad + (size_t)w * v.G + it.gslot;
    double* be = v.benergy + (size_t)w * v.B + it.eslot;
    if (it.type == BIAS_METAD) {
        if (lane == 0) {
            const GridBiasDev& g = it.grid;
            double bgv[MAX_BIAS_D];
            grid_eval(g, g.grid + (size_t)w * g.walkerStride, cv, be, bgv);
            for (int d = 0; d < g.D; d++) {
                bg[d] = bgv[d];
                if (bgL) bgL[it.gslot + d] = bgv[d];
            }
        }
    } else if (it.type == BIAS_OPES_EXPANDED) {
        const double invKT = it.p[0];
        const double* lw = st.stateConst + it.stateOff;
        const double* ci = lw + it.nState;  // CV value index of every state
        const double logZ = st.scal[(size_t)w * st.scalStride + 4 * i + 0];
        double mx = -1e300;
        for (int l = lane; l < it.nState; l += 64) mx = fmax(mx, lw[l] - cv[(int)ci[l]] * invKT);
        mx = wave_max1(mx);
        double se = 0.0;
        for (int l = lane; l < it.nState; l += 64) se += exp(lw[l] - cv[(int)ci[l]] * invKT - mx);
        se = wave_sum1(se);
        const double logQ = mx + log(se);
        if (lane == 0) *be = -(1.0 / invKT) * (logQ - logZ);
        for (int l = lane; l < it.nState; l += 64) {
            const double pl = exp(lw[l] - cv[(int)ci[l]] * invKT - logQ);
            bg[l] = pl;
            if (bgL) bgL[it.gslot + l] = pl;
        }
    } else if (it.type == BIAS_OPES_MULTITHERMAL) {
        const double kT0 = it.p[0];
        const double* db = st.stateConst + it.stateOff;
        const double* dF = st.stateVar + (size_t)w * st.stateVarStride + it.stateOff;
        const double U = cv[it.cv[0]];
        double mx = -1e300;
        for (int l = lane; l < it.nState; l += 64) mx = fmax(mx, -db[l] * U + dF[l] / kT0);
        mx = wave_max1(mx);
        double sd[2] = {0.0, 0.0};
        for (int l = lane; l < it.nState; l += 64) {
            const double e = exp(-db[l] * U + dF[l] / kT0 - mx);
            sd[0] += e;
            sd[1] += e * db[l];
        }
        wave_sum_n<2>(sd);
        if (lane == 0) {
            *be = -kT0 * (mx + log(sd[0] / (double)it.nState));
            const double g0 = kT0 * (sd[1] / sd[0]);
            bg[0] = g0;
            if (bgL) bgL[it.gslot] = g0;
        }
    } else if (EXTRA && lane == 0) {  // harmonic, linear, walls, PBMetaD
        if (it.type == BIAS_PBMETAD)
            pbmetad_eval(v, st, it, w, bgL);
        else
            restraints_eval(v, st, it, w, bgL);
    }
}

// one 64-thread workgroup (one wave) per (item, walker)
__global__ __launch_bounds__(64) void k_bias_small(StepView v, SmallBiasState st,
                                                   const SmallBiasItem* __restrict__ items, int n, int w0)
{
    small_bias_eval(v, st, items[blockIdx.x], blockIdx.x, w0 + blockIdx.y);
}

void launch_bias_small(const StepView& v, const SmallBiasState& st, const SmallBiasItem* items, int n, int w0, int nw,
                       hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bias_small, dim3(n, nw), dim3(64), 0, s, v, st, items, n, w0);
}

// ---------------------------------------------------------------------------
// Chunked sums: hill-list metadynamics (extension; truncation rule of
// CommonGluedKernels.cpp:3664-3677 applied at the CV point) and OPES
// (CommonGluedKernels.cpp:2642-2723).
//   hills: V = sum_k h_k exp(-E_k) [E_k <= 8], E_k = sum_d dc_d^2/(2 sigma_d^2)
//   OPES : prob = sum_k h_k max(G_k - kfloor, 0), acc_d = sum_k h_k G_k (-dc_d/s_kd^2)
// One 256-thread workgroup reduces one chunk of 1,024 entries (4 per thread,
// coalesced SoA loads) to D+1 FP64 partials.  Plain form: the workgroup that
// finishes last for a walker (atomic ticket) adds the partials and finalizes;
// batched form: a second launch does it (no cross-workgroup fences).
// ---------------------------------------------------------------------------
constexpr int BIAS_BLOCK = 256;
constexpr int PER_THREAD = 4;
constexpr int CHUNK = BIAS_BLOCK * PER_THREAD;
constexpr int HILL_BLOCK = BIAS_BLOCK;
constexpr int OPES_BLOCK = BIAS_BLOCK;
constexpr int PW = MAX_BIAS_D + 1;  // partial width in the fused buffer

int hill_chunks_for(int cap) { return (cap + CHUNK - 1) / CHUNK; }
int opes_chunks_for(int cap) { return (cap + CHUNK - 1) / CHUNK; }

// Block-wide sum of NV values, written by thread 0 to dst.  sh >= NV*4.
template <int NV>
__device__ inline void block_partial(double (&acc)[NV], double* sh, double* dst)
{
    wave_sum_n<NV>(acc);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int d = 0; d < NV; d++) sh[d * 4 + wid] = acc[d];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < NV; d++) dst[d] = sh[d * 4] + sh[d * 4 + 1] + sh[d * 4 + 2] + sh[d * 4 + 3];
    __syncthreads();
}

// One chunk of a hill list in two phases: the loads (every load of the chunk
// issued before any arithmetic) and the sum at the walker's CV values.
template <int D>
struct HillChunkRegs {
    double cx[PER_THREAD][D];
    double hk[PER_THREAD];  // 0 past the end of the list: such entries add exactly zero
};

// The loads go to indices clamped into the table's capacity, issued together
// with the load of the hill count: no load waits for the count.  Entries
// past the count are then set to zero (centre and height), so they add
// exactly zero as before.
template <int D>
__device__ inline void hill_chunk_load(const HillBiasDev& h, int w, int chunk, HillChunkRegs<D>& R)
{
    const int tb = h.shared ? 0 : w;
    const float* __restrict__ C = h.centers + (size_t)tb * D * h.cap;
    const float* __restrict__ H = h.heights + (size_t)tb * h.cap;
    const int n = h.count[tb];
    const int k0 = chunk * CHUNK + threadIdx.x;
#pragma unroll
    for (int r = 0; r < PER_THREAD; r++) {
        const int kc = min(k0 + r * BIAS_BLOCK, h.cap - 1);
#pragma unroll
        for (int d = 0; d < D; d++) R.cx[r][d] = (double)C[(size_t)d * h.cap + kc];
        R.hk[r] = (double)H[kc];
    }
#pragma unroll
    for (int r = 0; r < PER_THREAD; r++) {
        const bool ok = k0 + r * BIAS_BLOCK < n;
#pragma unroll
        for (int d = 0; d < D; d++) R.cx[r][d] = ok ? R.cx[r][d] : 0.0;
        R.hk[r] = ok ? R.hk[r] : 0.0;
    }
}

// The walker's CV values of a hill bias (scalar loads).
template <int D>
__device__ inline void hill_chunk_cv(const StepView& v, const HillBiasDev& h, int w, double (&s)[D])
{
    const double* cv = v.cvVal + (size_t)w * v.V;
#pragma unroll
    for (int d = 0; d < D; d++) s[d] = cv[h.cv[d]];
}

// exp(-e) for 0 <= e <= 8 from the FP32 exponential: e is rounded to FP32
// (an argument error of up to 4.8e-7 at e = 8, which the exponential would
// carry into the result) and the remainder d = e - e32, exact in FP64, is put
// back to first order, exp(-e) = exp(-e32) (1 - d) (second order < 1e-13).
// The relative error left is the FP32 exponential's own, ~1e-7.
__device__ inline double hill_exp(double e)
{
    const float ef = (float)e;
    const double d = e - (double)ef;
    return (double)expf(-ef) * (1.0 - d);
}

// The thread's sum over its hills at CV values s: V and dV/ds_d.  The
// exponent E is formed without contraction (its E <= 8 cut must fall where
// the oracle's does); the gradient terms use fused multiply-adds.  PER: some
// CV of the bias is periodic (a uniform branch picks the form, so that
// non-periodic biases skip the minimum-image arithmetic).
template <int D, bool PER>
__device__ inline void hill_terms_impl(const HillBiasDev& h, const double (&s)[D], const HillChunkRegs<D>& R,
                                       double* acc)
{
#pragma unroll
    for (int d = 0; d <= D; d++) acc[d] = 0.0;
#pragma unroll
    for (int r = 0; r < PER_THREAD; r++) {
        double e = 0.0, dc[D];
        {
#pragma clang fp contract(off)
#pragma unroll
            for (int d = 0; d < D; d++) {
                double x = s[d] - R.cx[r][d];
                if (PER && h.period[d] > 0.0) x = x - rint(x * h.invPeriod[d]) * h.period[d];
                dc[d] = x;
                e = e + x * x * h.i2s2[d];
            }
        }
        const double t = (e <= 8.0) ? R.hk[r] * hill_exp(e) : 0.0;
        acc[0] += t;
#pragma unroll
        for (int d = 0; d < D; d++) acc[1 + d] = fma(-t * dc[d], h.is2[d], acc[1 + d]);
    }
}

template <int D>
__device__ inline void hill_chunk_terms(const HillBiasDev& h, const double (&s)[D], const HillChunkRegs<D>& R,
                                        double* acc)
{
    bool per = false;
#pragma unroll
    for (int d = 0; d < D; d++) per = per || h.period[d] > 0.0;
    if (per)
        hill_terms_impl<D, true>(h, s, R, acc);
    else
        hill_terms_impl<D, false>(h, s, R, acc);
}

template <int D>
__device__ inline void hill_chunk_sum_at(const HillBiasDev& h, const double (&s)[D], const HillChunkRegs<D>& R,
                                         double* sh, double* dst)
{
    double acc[D + 1];
    hill_chunk_terms<D>(h, s, R, acc);
    block_partial<D + 1>(acc, sh, dst);
}

// A shared hill list: one workgroup evaluates its chunk for up to CHUNK_WB
// walkers (w0 ...), the chunk loaded once.  The walkers' sums go through one
// butterfly (each value on its own, as block_partial's) and the same
// wave-order sum, so every walker's partial equals a one-walker workgroup's
// bit for bit.  sh >= CHUNK_WB (D + 1) 4.
template <int D>
__device__ void hill_chunk_walkers(const StepView& v, const HillBiasDev& h, int w0, int nw, int chunk,
                                   int totalChunks, double* part, double* sh)
{
    constexpr int NV = CHUNK_WB * (D + 1);
    double s[CHUNK_WB][D];
#pragma unroll
    for (int j = 0; j < CHUNK_WB; j++) hill_chunk_cv<D>(v, h, w0 + min(j, nw - 1), s[j]);
    HillChunkRegs<D> R;
    hill_chunk_load<D>(h, w0, chunk, R);
    double acc[NV];
#pragma unroll
    for (int j = 0; j < CHUNK_WB; j++) hill_chunk_terms<D>(h, s[j], R, acc + j * (D + 1));
    wave_sum_n<NV>(acc);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k * 4 + wid] = acc[k];
    __syncthreads();
    if (threadIdx.x < nw) {
        const int j = threadIdx.x;
        double* dst = part + ((size_t)(w0 + j) * totalChunks + chunk) * PW;
#pragma unroll
        for (int d = 0; d <= D; d++) {
            const int k = j * (D + 1) + d;
            dst[d] = sh[k * 4] + sh[k * 4 + 1] + sh[k * 4 + 2] + sh[k * 4 + 3];
        }
    }
}

template <int D>
__device__ inline void hill_chunk_sum(const StepView& v, const HillBiasDev& h, int w, const HillChunkRegs<D>& R,
                                      double* sh, double* dst)
{
    double s[D];
    hill_chunk_cv<D>(v, h, w, s);
    hill_chunk_sum_at<D>(h, s, R, sh, dst);
}

// The CV values are requested before the hill loads: both chains are then
// in flight together (a wait for the scalar loads would otherwise also wait
// for the hill loads, which count on the same counter).
template <int D>
__device__ void hill_chunk(const StepView& v, const HillBiasDev& h, int w, int chunk, double* sh, double* dst)
{
    double s[D];
    hill_chunk_cv<D>(v, h, w, s);
    HillChunkRegs<D> R;
    hill_chunk_load<D>(h, w, chunk, R);
    hill_chunk_sum_at<D>(h, s, R, sh, dst);
}

template <int D>
__device__ void opes_chunk(const StepView& v, const OpesBiasDev& o, int w, int chunk, double* sh, double* dst)
{
    const int tb = o.shared ? 0 : w;
    const double* __restrict__ C = o.centers + (size_t)tb * D * o.cap;
    const double* __restrict__ Sg = o.sigmas + (size_t)tb * D * o.cap;
    const double* __restrict__ L = o.logw + (size_t)tb * o.cap;
    const int nk = o.ints[tb * 4 + 0];
    const double* cv = v.cvVal + (size_t)w * v.V;
    double s[D];
#pragma unroll
    for (int d = 0; d < D; d++) s[d] = cv[o.cv[d]];
    const double kfloor = exp(-0.5 * o.cutoff2);
    double acc[D + 1];
#pragma unroll
    for (int d = 0; d <= D; d++) acc[d] = 0.0;
    for (int r = 0; r < PER_THREAD; r++) {
        const int k = chunk * CHUNK + r * BIAS_BLOCK + threadIdx.x;
        if (k >= nk) break;
        double n2 = 0.0, dc[D], sg[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
#pragma clang fp contract(off)
            dc[d] = s[d] - C[(size_t)d * o.cap + k];
            sg[d] = Sg[(size_t)d * o.cap + k];
            n2 = n2 + dc[d] * dc[d] / (sg[d] * sg[d]);
        }
        if (!(n2 < o.cutoff2)) continue;
        const double G = exp(-0.5 * n2), kern = G - kfloor;
        if (kern <= 0.0) continue;
        const double hk = exp(L[k]);
        if (!(hk > 0.0 && hk < 1e308)) continue;
        acc[0] += hk * kern;
#pragma unroll
        for (int d = 0; d < D; d++) acc[1 + d] += hk * G * (-dc[d] / (sg[d] * sg[d]));
    }
    block_partial<D + 1>(acc, sh, dst);
}

// Sum of n partial vectors (width wd <= PW, stride ld) by wave 0 of the
// workgroup: lane l adds entries l, l + 64, ... in index order, then a
// butterfly over the wave (commutative pairwise adds: every lane ends with the
// same bits).  The order depends only on n, so the plain and the batched forms
// agree bit for bit.  ATOMIC: the partials were written by other workgroups of
// the same launch (read with device-scope loads after the ticket).
template <bool ATOMIC, int WD>
__device__ inline void reduce_partials_w(const double* src, int n, int ld, double* tot)
{
    const int lane = threadIdx.x & 63;
    double a[WD];
#pragma unroll
    for (int d = 0; d < WD; d++) a[d] = 0.0;
    for (int j = lane; j < n; j += 64) {
#pragma unroll
        for (int d = 0; d < WD; d++) {
            const double* p = src + (size_t)j * ld + d;
            a[d] += ATOMIC ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
        }
    }
    wave_sum_n<WD>(a);
#pragma unroll
    for (int d = 0; d < PW; d++) tot[d] = d < WD ? a[d < WD ? d : 0] : 0.0;
}

// The butterfly covers exactly the wd = D + 1 values of the bias.
template <bool ATOMIC>
__device__ inline void reduce_partials(const double* src, int n, int wd, int ld, double* tot)
{
    switch (wd) {
        case 1: reduce_partials_w<ATOMIC, 1>(src, n, ld, tot); break;
        case 2: reduce_partials_w<ATOMIC, 2>(src, n, ld, tot); break;
        case 3: reduce_partials_w<ATOMIC, 3>(src, n, ld, tot); break;
        case 4: reduce_partials_w<ATOMIC, 4>(src, n, ld, tot); break;
        case 5: reduce_partials_w<ATOMIC, 5>(src, n, ld, tot); break;
        case 6: reduce_partials_w<ATOMIC, 6>(src, n, ld, tot); break;
        case 7: reduce_partials_w<ATOMIC, 7>(src, n, ld, tot); break;
        case 8: reduce_partials_w<ATOMIC, 8>(src, n, ld, tot); break;
        default: reduce_partials_w<ATOMIC, PW>(src, n, ld, tot); break;
    }
}

__device__ inline void hill_finalize(const StepView& v, const HillBiasDev& h, int w, const double* tot,
                                     double* bgL = nullptr)
{
    v.benergy[(size_t)w * v.B + h.eslot] = tot[0];
    for (int d = 0; d < h.D; d++) {
        v.bgrad[(size_t)w * v.G + h.gslot + d] = tot[1 + d];
        if (bgL) bgL[h.gslot + d] = tot[1 + d];
    }
}

__device__ inline void opes_finalize(const StepView& v, const OpesBiasDev& o, int w, const double* tot,
                                     double* bgL = nullptr)
{
    const int tb = o.shared ? 0 : w;
    const int nk = o.ints[tb * 4 + 0];
    double* be = v.benergy + (size_t)w * v.B + o.eslot;
    double* bg = v.bgrad + (size_t)w * v.G + o.gslot;
    if (nk == 0) {
        *be = -o.barrier;
        for (int d = 0; d < o.D; d++) {
            bg[d] = 0.0;
            if (bgL) bgL[o.gslot + d] = 0.0;
        }
        return;
    }
    const double prob = tot[0];
    const double su = o.dbl[(size_t)tb * (4 + 2 * o.D) + 0];
    double pz = su > 0.0 ? prob * (double)nk / su : 0.0;
    if (!(pz >= 0.0)) pz = 0.0;
    const double arg = pz + o.eps;
    *be = o.invGF * o.kT * log(arg > 1e-30 ? arg : 1e-30);
    const double f = prob > 0.0 ? o.invGF * o.kT * (pz / (pz + o.eps)) / prob : 0.0;
    for (int d = 0; d < o.D; d++) {
        bg[d] = f * tot[1 + d];
        if (bgL) bgL[o.gslot + d] = f * tot[1 + d];
    }
}

// ---- plain form: one launch per chunked bias ----
template <int D>
__global__ __launch_bounds__(BIAS_BLOCK) void k_bias_hills(StepView v, HillBiasDev h, int w0)
{
    __shared__ double sh[(D + 1) * 4];
    const int w = w0 + blockIdx.y;
    hill_chunk<D>(v, h, w, blockIdx.x, sh, h.part + ((size_t)w * h.chunks + blockIdx.x) * (D + 1));
    if (!last_block(&h.counter[w], gridDim.x)) return;
    if (threadIdx.x >= 64) return;
    double tot[PW];
    reduce_partials<true>(h.part + (size_t)w * h.chunks * (D + 1), gridDim.x, D + 1, D + 1, tot);
    if (threadIdx.x == 0) {
        hill_finalize(v, h, w, tot);
        h.counter[w] = 0u;
    }
}

template <int D>
__global__ __launch_bounds__(BIAS_BLOCK) void k_bias_opes(StepView v, OpesBiasDev o, int w0)
{
    __shared__ double sh[(D + 1) * 4];
    const int w = w0 + blockIdx.y;
    opes_chunk<D>(v, o, w, blockIdx.x, sh, o.part + ((size_t)w * o.chunks + blockIdx.x) * (D + 1));
    if (!last_block(&o.counter[w], gridDim.x)) return;
    if (threadIdx.x >= 64) return;
    double tot[PW];
    reduce_partials<true>(o.part + (size_t)w * o.chunks * (D + 1), gridDim.x, D + 1, D + 1, tot);
    if (threadIdx.x == 0) {
        opes_finalize(v, o, w, tot);
        o.counter[w] = 0u;
    }
}

void launch_bias_hills(const StepView& v, const HillBiasDev& h, int w0, int nw, hipStream_t s)
{
    dim3 g(h.chunks, nw), b(BIAS_BLOCK);
    switch (h.D) {
        case 1: hipLaunchKernelGGL(k_bias_hills<1>, g, b, 0, s, v, h, w0); break;
        case 2: hipLaunchKernelGGL(k_bias_hills<2>, g, b, 0, s, v, h, w0); break;
        case 3: hipLaunchKernelGGL(k_bias_hills<3>, g, b, 0, s, v, h, w0); break;
        default: hipLaunchKernelGGL(k_bias_hills<4>, g, b, 0, s, v, h, w0); break;
    }
}

void launch_bias_opes(const StepView& v, const OpesBiasDev& o, int w0, int nw, hipStream_t s)
{
    dim3 g(o.chunks, nw), b(BIAS_BLOCK);
    switch (o.D) {
        case 1: hipLaunchKernelGGL(k_bias_opes<1>, g, b, 0, s, v, o, w0); break;
        case 2: hipLaunchKernelGGL(k_bias_opes<2>, g, b, 0, s, v, o, w0); break;
        case 3: hipLaunchKernelGGL(k_bias_opes<3>, g, b, 0, s, v, o, w0); break;
        default: hipLaunchKernelGGL(k_bias_opes<4>, g, b, 0, s, v, o, w0); break;
    }
}

// ---------------------------------------------------------------------------
// Batched form, two launches for the whole bias + scatter stage: every chunk
// of every chunked bias of every walker, then one workgroup per walker that
// finalizes the chunked biases, evaluates the one-thread biases, forms
// dU/dCV and (when the walker's atom list is short) performs the force
// scatter.  The launch boundary orders the partials (no cross-workgroup
// fences).  Chunking and summation order are those of the plain kernels, so
// both forms give bit-identical results.
// ---------------------------------------------------------------------------
constexpr int MAX_LDS_VALUES = 1024;

// KIND 0: any mix of hill and OPES biases of any D; KIND 1 (2): only hill
// (OPES) biases, all of dimension DD -- the specialisations keep register use
// low enough for every workgroup of a typical step to be resident at once.
template <int KIND, int DD>
__device__ inline void chunk_block(const StepView& v, const FusedBiasPlan& fp, int b, int w, double* sh)
{
    if (KIND == 1 && fp.nChunked == 1) {  // one hill bias: its descriptor from the kernel arguments
        hill_chunk<DD>(v, fp.hill0, w, b, sh, fp.part + ((size_t)w * fp.totalChunks + b) * PW);
        return;
    }
    int k = 0;
    while (k + 1 < fp.nChunked && b >= fp.chunkStart[k + 1]) k++;
    const int chunk = b - fp.chunkStart[k];
    double* dst = fp.part + ((size_t)w * fp.totalChunks + b) * PW;
    if (KIND == 1) {
        hill_chunk<DD>(v, fp.hills[k], w, chunk, sh, dst);
    } else if (KIND == 2) {
        opes_chunk<DD>(v, fp.opes[k], w, chunk, sh, dst);
    } else if (k < fp.nHill) {
        const HillBiasDev& h = fp.hills[k];
        switch (h.D) {
            case 1: hill_chunk<1>(v, h, w, chunk, sh, dst); break;
            case 2: hill_chunk<2>(v, h, w, chunk, sh, dst); break;
            case 3: hill_chunk<3>(v, h, w, chunk, sh, dst); break;
            default: hill_chunk<4>(v, h, w, chunk, sh, dst); break;
        }
    } else {
        const OpesBiasDev& o = fp.opes[k - fp.nHill];
        switch (o.D) {
            case 1: opes_chunk<1>(v, o, w, chunk, sh, dst); break;
            case 2: opes_chunk<2>(v, o, w, chunk, sh, dst); break;
            case 3: opes_chunk<3>(v, o, w, chunk, sh, dst); break;
            default: opes_chunk<4>(v, o, w, chunk, sh, dst); break;
        }
    }
}

template <int KIND, int DD>
__global__ __launch_bounds__(BIAS_BLOCK) void k_step_bias_chunks(StepView v, FusedBiasPlan fp)
{
    __shared__ double sh[PW * 4];
    chunk_block<KIND, DD>(v, fp, blockIdx.x, blockIdx.y, sh);
}

// One shared hill list: CHUNK_WB walkers per workgroup (a kernel of its own,
// so that the one-walker chunk kernel keeps its small register budget).
template <int DD>
__global__ __launch_bounds__(BIAS_BLOCK) void k_step_hill_chunks_shared(StepView v, FusedBiasPlan fp)
{
    __shared__ double sh[CHUNK_WB * (DD + 1) * 4];
    const int w0 = blockIdx.y * CHUNK_WB;
    hill_chunk_walkers<DD>(v, fp.hill0, w0, min(CHUNK_WB, fp.nWalkers - w0), blockIdx.x, fp.totalChunks, fp.part,
                           sh);
}

static void launch_chunks(const StepView& v, const FusedBiasPlan& fpIn, int nw, hipStream_t s)
{
    FusedBiasPlan fp = fpIn;
    fp.nWalkers = nw;
    if (fp.kind == 1 && fp.walkersPerChunk > 1 && fp.kindD >= 1 && fp.kindD <= 3) {
        const dim3 g(fp.totalChunks, (nw + CHUNK_WB - 1) / CHUNK_WB), b(BIAS_BLOCK);
        if (fp.kindD == 1) hipLaunchKernelGGL(k_step_hill_chunks_shared<1>, g, b, 0, s, v, fp);
        else if (fp.kindD == 2) hipLaunchKernelGGL(k_step_hill_chunks_shared<2>, g, b, 0, s, v, fp);
        else hipLaunchKernelGGL(k_step_hill_chunks_shared<3>, g, b, 0, s, v, fp);
        return;
    }
    const dim3 g(fp.totalChunks, nw), b(BIAS_BLOCK);
    if (fp.kind == 1 && fp.kindD == 1) hipLaunchKernelGGL((k_step_bias_chunks<1, 1>), g, b, 0, s, v, fp);
    else if (fp.kind == 1 && fp.kindD == 2) hipLaunchKernelGGL((k_step_bias_chunks<1, 2>), g, b, 0, s, v, fp);
    else if (fp.kind == 1 && fp.kindD == 3) hipLaunchKernelGGL((k_step_bias_chunks<1, 3>), g, b, 0, s, v, fp);
    else if (fp.kind == 2 && fp.kindD == 1) hipLaunchKernelGGL((k_step_bias_chunks<2, 1>), g, b, 0, s, v, fp);
    else if (fp.kind == 2 && fp.kindD == 2) hipLaunchKernelGGL((k_step_bias_chunks<2, 2>), g, b, 0, s, v, fp);
    else if (fp.kind == 2 && fp.kindD == 3) hipLaunchKernelGGL((k_step_bias_chunks<2, 3>), g, b, 0, s, v, fp);
    else hipLaunchKernelGGL((k_step_bias_chunks<0, 1>), g, b, 0, s, v, fp);
}

// One workgroup per walker; each wave has one chain of dependent loads, so
// the slowest chain, not their sum, sets the time to the barrier:
//   wave 0     reduces the chunk partials (written by the previous launch)
//              and finalizes the chunked biases;
//   wave 1     evaluates the one-thread biases (item, CV values, grid
//              corners: three dependent loads);
//   waves 2-3  copy the dU/dCV combine table into LDS.
// Waves 0, 2 and 3 also load their scatter slots into registers (row range,
// first row's CV and Jacobian, GPU atom index).  Gradient slots go to global
// memory and to LDS.  After one barrier dU/dCV is formed from LDS and the
// scatter needs only arithmetic and the fixed-point atomics.
constexpr int FIN_SLOTS = 3;           // scatter slots per slot-owning thread held in registers
constexpr int FIN_OWNERS = 3 * 64;     // slot-owning threads (waves 0, 2, 3)
constexpr int FIN_LDS_G = 512;         // bias-gradient slots held in LDS
constexpr int FIN_LDS_COMB = 1024;     // combine entries held in LDS

// Returns the LDS copy of walker w's dU/dCV (the first MAX_LDS_VALUES
// values; the rest are in global dUdCV, written by this workgroup).
template <bool EXTRA>
__device__ const double* finalize_walker(const StepView& v, const FusedBiasPlan& fp, int w)
{
    __shared__ double sU[MAX_LDS_VALUES];
    __shared__ double sBg[FIN_LDS_G];
    __shared__ int sCombStart[MAX_LDS_VALUES + 1];
    __shared__ int sCombSlot[FIN_LDS_COMB];
    const int tid = threadIdx.x;
    const ScatterPlan& sp = fp.sp;
    const bool ldsComb = v.V <= MAX_LDS_VALUES && v.G <= FIN_LDS_G && sp.nComb <= FIN_LDS_COMB && !sp.testGrad;
    double* bgL = ldsComb ? sBg : nullptr;
    // bias-independent loads
    const int wave = tid >> 6;
    const int owner = wave == 0 ? tid : tid - 64;  // slot owner index (waves 0, 2, 3)
    const bool doScatter = fp.scatter != 0 && v.S > 0 && v.R > 0;
    const bool ownsSlots = doScatter && wave != 1;
    int sb[FIN_SLOTS], se[FIN_SLOTS], sc0[FIN_SLOTS], sg[FIN_SLOTS];
    float4 sj[FIN_SLOTS];
    const size_t rbase = (size_t)w * v.R;
    if (ownsSlots) {
#pragma unroll
        for (int r = 0; r < FIN_SLOTS; r++) {  // clamped slot: no divergent loads
            const int s = min(owner + r * FIN_OWNERS, v.S - 1);
            sb[r] = sp.segStart[s];
            se[r] = sp.segStart[s + 1];
            sg[r] = v.gidx[(size_t)w * v.S + s];
        }
#pragma unroll
        for (int r = 0; r < FIN_SLOTS; r++) {
            const int k = min(sb[r], v.R - 1);  // a slot without rows reads a valid row and drops it
            sc0[r] = sp.entryCv[k];
            sj[r] = v.jrow[rbase + k];
        }
    }
    if (ldsComb && wave >= 2) {
        for (int c = tid - 128; c <= v.V; c += 128) sCombStart[c] = sp.combStart[c];
        for (int k = tid - 128; k < sp.nComb; k += 128) sCombSlot[k] = sp.combSlot[k];
    }
    if (wave == 0) {
        for (int k = 0; k < fp.nChunked; k++) {
            const int c0 = fp.chunkStart[k], nc = fp.chunkStart[k + 1] - c0;
            const double* src = fp.part + ((size_t)w * fp.totalChunks + c0) * PW;
            const int D = fp.kind ? fp.kindD : (k < fp.nHill ? fp.hills[k].D : fp.opes[k - fp.nHill].D);
            double tot[PW];
            reduce_partials<false>(src, nc, D + 1, PW, tot);
            if (tid == 0) {
                if (k < fp.nHill)
                    hill_finalize(v, fp.hills[k], w, tot, bgL);
                else
                    opes_finalize(v, fp.opes[k - fp.nHill], w, tot, bgL);
            }
        }
    } else if (wave == 1) {
        for (int i = 0; i < fp.nSmall; i++) small_bias_eval<EXTRA>(v, fp.st, fp.small[i], i, w, bgL);
    }
    __syncthreads();
    if (!ldsComb) {
        if (doScatter)
            scatter_walker_block(v, sp, w, sU, MAX_LDS_VALUES);
        else
            combine_dudcv(v, sp, w, sU, MAX_LDS_VALUES, true);
        return sU;
    }
    // dU/dCV from LDS (the combine order of combine_dudcv)
    for (int c = tid; c < v.V; c += BIAS_BLOCK) {
        double u = 0.0;
        for (int k = sCombStart[c]; k < sCombStart[c + 1]; k++) u += sBg[sCombSlot[k]];
        sU[c] = u;
        v.dUdCV[(size_t)w * v.V + c] = u;
    }
    __syncthreads();
    if (!doScatter) return sU;
    const WalkerInfo wi = v.walkers[w];
    long long* F = v.force + wi.forceOff;
    if (ownsSlots) {
#pragma unroll
        for (int r = 0; r < FIN_SLOTS; r++) {
            const int s = owner + r * FIN_OWNERS;
            if (s >= v.S) break;
            // first row from registers, further rows of this atom (rare) from memory,
            // in the order of scatter_slot
            if (sb[r] == se[r]) continue;
            const double u0 = sU[sc0[r]];
            double fx = -(u0 * (double)sj[r].x), fy = -(u0 * (double)sj[r].y), fz = -(u0 * (double)sj[r].z);
            for (int k = sb[r] + 1; k < se[r]; k++) {
                const double u = sU[sp.entryCv[k]];
                const float4 j = v.jrow[rbase + k];
                fx -= u * (double)j.x;
                fy -= u * (double)j.y;
                fz -= u * (double)j.z;
            }
            atomic_add_fixed(F + sg[r], fx);
            atomic_add_fixed(F + sg[r] + wi.padded, fy);
            atomic_add_fixed(F + sg[r] + 2 * (size_t)wi.padded, fz);
        }
    }
    for (int s = tid + FIN_SLOTS * FIN_OWNERS; s < v.S; s += BIAS_BLOCK) scatter_slot(v, sp, w, s, sU, MAX_LDS_VALUES);
    return sU;
}

template <bool EXTRA>
__global__ __launch_bounds__(BIAS_BLOCK) void k_step_bias_finalize(StepView v, FusedBiasPlan fp)
{
    finalize_walker<EXTRA>(v, fp, blockIdx.x);
}

// The finalize and the slot scatter in one launch, for plans whose atom slots
// are too many for one workgroup per walker: every workgroup of walker
// blockIdx.y forms dU/dCV itself (the finalize's code, so the same bits; the
// global outputs are written with the same values by each of them), then
// scatters its BIAS_BLOCK slots.  Saves the scatter launch and its ramp.
template <bool EXTRA>
__global__ __launch_bounds__(BIAS_BLOCK) void k_step_finalize_scatter(StepView v, FusedBiasPlan fp)
{
    const int w = blockIdx.y;
    const double* sU = finalize_walker<EXTRA>(v, fp, w);
    const int s = blockIdx.x * BIAS_BLOCK + threadIdx.x;
    if (s < v.S) scatter_slot(v, fp.sp, w, s, sU, MAX_LDS_VALUES);
}

// ---------------------------------------------------------------------------
// The whole batched step of a small plan in one workgroup per walker:
// few-atom CVs, then every chunk of the chunked biases in index order (the
// chunk workgroups' code, partials to the same buffer), then the finalize.
// One launch instead of three for plans whose launches are mostly latency
// (a few atoms, a hill list of a few thousand entries); the same device
// code in the same order, so the results equal the other forms bit for bit.
// Every hand-off between waves stays inside the workgroup (barriers).
// ---------------------------------------------------------------------------
constexpr int TINY_PREFETCH_CHUNKS = 4;  // one hill list of up to 4,096 entries: loads issued before the CVs

template <int KIND, int DD, bool EXTRA = false>
__global__ __launch_bounds__(BIAS_BLOCK) void k_step_tiny(StepView v, const SmallCVItem* __restrict__ cvItems, int nCv,
                                                          FusedBiasPlan fp)
{
    __shared__ double sh[PW * 4];
    const int w = blockIdx.x;
    if constexpr (KIND == 1) {
        if (fp.nChunked == 1 && fp.totalChunks <= TINY_PREFETCH_CHUNKS) {
            // the hill loads do not depend on the CVs: in flight during the CV evaluation
            const HillBiasDev& h = fp.hill0;
            HillChunkRegs<DD> R[TINY_PREFETCH_CHUNKS];
#pragma unroll
            for (int c = 0; c < TINY_PREFETCH_CHUNKS; c++)
                if (c < fp.totalChunks) hill_chunk_load<DD>(h, w, c, R[c]);
            for (int i = threadIdx.x; i < nCv; i += BIAS_BLOCK) small_item(v, cvItems[i], w);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < TINY_PREFETCH_CHUNKS; c++)
                if (c < fp.totalChunks)
                    hill_chunk_sum<DD>(v, h, w, R[c], sh, fp.part + ((size_t)w * fp.totalChunks + c) * PW);
            finalize_walker<EXTRA>(v, fp, w);
            return;
        }
    }
    for (int i = threadIdx.x; i < nCv; i += BIAS_BLOCK) small_item(v, cvItems[i], w);
    __syncthreads();
    for (int b = 0; b < fp.totalChunks; b++) chunk_block<KIND, DD>(v, fp, b, w, sh);  // each ends with a barrier
    finalize_walker<EXTRA>(v, fp, w);
}

void launch_step_tiny(const StepView& v, const SmallCVItem* cvItems, int nCv, const FusedBiasPlan& fp, int nw,
                      hipStream_t s)
{
    const dim3 g(nw), b(BIAS_BLOCK);
    if (fp.extra) hipLaunchKernelGGL((k_step_tiny<0, 1, true>), g, b, 0, s, v, cvItems, nCv, fp);
    else if (fp.kind == 1 && fp.kindD == 1) hipLaunchKernelGGL((k_step_tiny<1, 1>), g, b, 0, s, v, cvItems, nCv, fp);
    else if (fp.kind == 1 && fp.kindD == 2) hipLaunchKernelGGL((k_step_tiny<1, 2>), g, b, 0, s, v, cvItems, nCv, fp);
    else if (fp.kind == 1 && fp.kindD == 3) hipLaunchKernelGGL((k_step_tiny<1, 3>), g, b, 0, s, v, cvItems, nCv, fp);
    else if (fp.kind == 2 && fp.kindD == 1) hipLaunchKernelGGL((k_step_tiny<2, 1>), g, b, 0, s, v, cvItems, nCv, fp);
    else if (fp.kind == 2 && fp.kindD == 2) hipLaunchKernelGGL((k_step_tiny<2, 2>), g, b, 0, s, v, cvItems, nCv, fp);
    else if (fp.kind == 2 && fp.kindD == 3) hipLaunchKernelGGL((k_step_tiny<2, 3>), g, b, 0, s, v, cvItems, nCv, fp);
    else hipLaunchKernelGGL((k_step_tiny<0, 1>), g, b, 0, s, v, cvItems, nCv, fp);
}

void launch_step_bias_chunks(const StepView& v, const FusedBiasPlan& fp, int nw, hipStream_t s)
{
    launch_chunks(v, fp, nw, s);
}

void launch_step_bias(const StepView& v, const FusedBiasPlan& fp, int nw, hipStream_t s)
{
    if (fp.totalChunks > 0) launch_chunks(v, fp, nw, s);
    if (fp.finalizeScatter && !fp.scatter) {
        const dim3 g((v.S + BIAS_BLOCK - 1) / BIAS_BLOCK, nw);
        if (fp.extra)
            hipLaunchKernelGGL(k_step_finalize_scatter<true>, g, dim3(BIAS_BLOCK), 0, s, v, fp);
        else
            hipLaunchKernelGGL(k_step_finalize_scatter<false>, g, dim3(BIAS_BLOCK), 0, s, v, fp);
        return;
    }
    if (fp.extra)
        hipLaunchKernelGGL(k_step_bias_finalize<true>, dim3(nw), dim3(BIAS_BLOCK), 0, s, v, fp);
    else
        hipLaunchKernelGGL(k_step_bias_finalize<false>, dim3(nw), dim3(BIAS_BLOCK), 0, s, v, fp);
}

// ---------------------------------------------------------------------------
// Deposition.
// ---------------------------------------------------------------------------

// Grid deposit, CommonGluedKernels.cpp:3645-3679.  A workgroup covers 64
// grid nodes; the walkers' heights h0 exp(-V hif) and centres are formed once
// per workgroup in LDS, and only the walkers whose Gaussian can reach one of
// the workgroup's nodes are kept (in walker order): a walker is dropped when,
// in some dimension, its centre lies farther than 4 sigma (plus a 1e-6
// relative margin) from the range of the nodes' coordinates, every image of
// a periodic centre included.  The four waves take every fourth kept walker
// of a shared grid (a private grid is one y-block row per walker) and the
// four partial sums of a node are added in wave order, then once to the grid
// (deterministic, no atomics).  A node is skipped for a walker when some
// |dc_d| > 4 sigma_d (with a 1e-9 relative margin), which implies E > 8;
// otherwise E is formed with the reference's expression, unfused, so the
// E <= 8 predicate is the oracle's.
constexpr int DEP_BLOCK = 512;
constexpr int DEP_NODES = 64;  // one wave: wave 0 forms the nodes' coordinate ranges
constexpr int DEP_GROUPS = DEP_BLOCK / DEP_NODES;
constexpr int DEP_WALKERS = DEP_BLOCK;  // walkers staged in LDS per pass (one per thread)

__global__ __launch_bounds__(DEP_BLOCK) void k_dep_grid(StepView v, GridBiasDev g, int w0, int nw)
{
#pragma clang fp contract(off)
    __shared__ double sC[DEP_WALKERS][3];
    __shared__ double sH[DEP_WALKERS];
    __shared__ double sAdd[DEP_GROUPS][DEP_NODES];
    __shared__ double sLo[3], sHi[3];
    __shared__ int sKeep[DEP_BLOCK / 64];
    const int nl = threadIdx.x % DEP_NODES, grp = threadIdx.x / DEP_NODES;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int idx = blockIdx.x * DEP_NODES + nl;
    const bool priv = g.walkerStride != 0;
    const int wFirst = priv ? w0 + blockIdx.y : w0;
    const int wCount = priv ? 1 : nw;
    // per node and dimension: position, period, 2 sigma^2 (the divisions
    // dc / period and dc^2 / (2 sigma^2) stay divisions, as in the reference
    // and the oracle, so that the E <= 8 predicate is theirs bit for bit)
    double pos[3] = {0.0, 0.0, 0.0}, per[3] = {0.0, 0.0, 0.0}, two_s2[3], far2[3];
    for (int d = 0; d < g.D; d++) {
        const int bi = (min(idx, g.total - 1) / g.strides[d]) % g.npts[d];
        pos[d] = g.origin[d] + (double)bi / g.invSpacing[d];
        if (g.periodic[d]) per[d] = (double)g.npts[d] / g.invSpacing[d];
        two_s2[d] = 2.0 * g.sigma[d] * g.sigma[d];
        far2[d] = 16.0 * g.sigma[d] * g.sigma[d] * (1.0 + 1e-9);
    }
    if (wid == 0) {  // the range of the workgroup's node coordinates per dimension
        for (int d = 0; d < g.D; d++) {
            double lo = pos[d], hi = pos[d];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                lo = fmin(lo, __shfl_xor(lo, o, 64));
                hi = fmax(hi, __shfl_xor(hi, o, 64));
            }
            if (lane == 0) {
                sLo[d] = lo;
                sHi[d] = hi;
            }
        }
    }
    // the node's grid value, requested now (its round trip overlaps the work below)
    double* grid = g.grid + (size_t)wFirst * g.walkerStride;
    const double old = grp == 0 ? grid[min(idx, g.total - 1)] : 0.0;
    double add = 0.0;
    for (int base = 0; base < wCount; base += DEP_WALKERS) {
        const int m = min(DEP_WALKERS, wCount - base);
        __syncthreads();  // sLo / sHi written; the previous pass done with sC / sH
        // thread t: walker base + t, kept when its Gaussian can reach a node
        const int t = threadIdx.x;
        double c[3] = {0.0, 0.0, 0.0}, h = 0.0;
        bool keep = t < m;
        if (keep) {
            const int w = wFirst + base + t;
            const double* cv = v.cvVal + (size_t)w * v.V;
            const double V = v.benergy[(size_t)w * v.B + g.eslot];
            h = g.hif > 0.0 ? g.h0 * exp(-V * g.hif) : g.h0;
            for (int d = 0; d < g.D; d++) {
                c[d] = cv[g.cv[d]];
                const double lo = sLo[d], hi = sHi[d];
                auto gap = [&](double x) { return fmax(0.0, fmax(lo - x, x - hi)); };
                double dist = gap(c[d]);
                if (g.periodic[d]) {
                    // the centre's image nearest the range's midpoint and its two
                    // neighbours: a node of the range is within one period of it
                    const double L = per[d];
                    const double cc = c[d] - L * rint((c[d] - 0.5 * (lo + hi)) / L);
                    dist = fmin(gap(cc), fmin(gap(cc + L), gap(cc - L)));
                }
                if (dist * dist > far2[d] * (1.0 + 1e-6)) keep = false;
            }
        }
        // compact the kept walkers in walker order
        const unsigned long long bal = __ballot(keep);
        if (lane == 0) sKeep[wid] = __popcll(bal);
        __syncthreads();
        int off = 0, mk = 0;
#pragma unroll
        for (int q = 0; q < DEP_BLOCK / 64; q++) {
            off += q < wid ? sKeep[q] : 0;
            mk += sKeep[q];
        }
        if (keep) {
            const int k = off + (int)__popcll(bal & ((1ull << lane) - 1ull));
            sH[k] = h;
            for (int d = 0; d < g.D; d++) sC[k][d] = c[d];
        }
        __syncthreads();
        for (int u = grp; u < mk; u += DEP_GROUPS) {
            double e = 0.0;
            bool near = true;
            for (int d = 0; d < g.D && near; d++) {
                double dc = pos[d] - sC[u][d];
                if (g.periodic[d]) dc -= round(dc / per[d]) * per[d];
                if (dc * dc > far2[d]) {
                    near = false;
                } else {
                    e += dc * dc / two_s2[d];
                }
            }
            if (near && e <= 8.0) add += sH[u] * exp(-e);
        }
    }
    sAdd[grp][nl] = add;
    __syncthreads();
    if (grp == 0 && idx < g.total) {
        double tot = sAdd[0][nl];
#pragma unroll
        for (int q = 1; q < DEP_GROUPS; q++) tot += sAdd[q][nl];
        grid[idx] = old + tot;
    }
}

void launch_dep_grid(const StepView& v, const GridBiasDev& g, int w0, int nw, hipStream_t s)
{
    dim3 grid((g.total + DEP_NODES - 1) / DEP_NODES, g.walkerStride ? nw : 1);
    hipLaunchKernelGGL(k_dep_grid, grid, dim3(DEP_BLOCK), 0, s, v, g, w0, nw);
}

// PBMetaD deposit, CommonGluedKernels.cpp:3766-3795 and :5147-5154: every
// sub-grid i takes a Gaussian at s_i of height w_i h0 exp(-V_i hif) (w_i,
// V_i of the last evaluation).  One workgroup per (sub-grid, table), a thread
// per node; a shared table takes the walkers in walker order, each node
// adding them in that order.
__global__ void k_dep_pb(StepView v, SmallBiasState st, SmallBiasItem it, int w0, int nw)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x;
    const bool shared = it.p[3] != 0.0;
    const int wFirst = shared ? w0 : w0 + blockIdx.y, wLast = shared ? w0 + nw : wFirst + 1;
    const int tb = shared ? 0 : wFirst;
    const double* c = st.stateConst + it.stateOff + PB_CONST * i;
    const int np = (int)c[2], N = it.D;
    double* grid = st.pbGrid + it.bias + (size_t)tb * it.nState + (int)c[5];
    const double per = (double)np / c[1], sigma = c[4];
    for (int idx = threadIdx.x; idx < np; idx += blockDim.x) {
        const double pos = c[0] + (double)idx / c[1];
        double acc = grid[idx];
        for (int w = wFirst; w < wLast; w++) {
            const double* var = st.stateVar + (size_t)w * st.stateVarStride + it.stateOff;
            const double Vi = var[i], wi = var[N + i];
            const double height = it.p[2] > 0.0 ? wi * it.p[0] * exp(-Vi * it.p[2]) : wi * it.p[0];
            double dc = pos - v.cvVal[(size_t)w * v.V + it.cv[i]];
            if (c[3] != 0.0) dc -= round(dc / per) * per;
            const double e = dc * dc / (2.0 * sigma * sigma);
            if (e <= 8.0) acc += height * exp(-e);
        }
        grid[idx] = acc;
    }
}

void launch_dep_pb(const StepView& v, const SmallBiasState& st, const SmallBiasItem& it, int w0, int nw,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_dep_pb, dim3(it.D, it.p[3] != 0.0 ? 1 : nw), dim3(256), 0, s, v, st, it, w0, nw);
}

// Hill deposit: centre = current CVs, height = h0 exp(-V hif) (the WT rule of
// CommonGluedKernels.cpp:3662 with this bias's own V).  Private lists: one
// thread per walker.  A shared list takes the walkers in walker order, as the
// reference's sequential deposits do: walker w0 + i goes to slot count + i,
// and the walkers past the capacity are dropped.  A height depends only on
// that walker's V of the last evaluation, so the appends are independent; one
// workgroup reads the count before any thread advances it.
constexpr int DEP_HILLS_SHARED_BLOCK = 256;

__device__ inline void deposit_hill(const StepView& v, const HillBiasDev& h, int tb, int w, int slot)
{
    float* C = h.centers + (size_t)tb * h.D * h.cap;
    float* H = h.heights + (size_t)tb * h.cap;
    const double* cv = v.cvVal + (size_t)w * v.V;
    const double V = v.benergy[(size_t)w * v.B + h.eslot];
    for (int d = 0; d < h.D; d++) C[(size_t)d * h.cap + slot] = (float)cv[h.cv[d]];
    H[slot] = (float)(h.hif > 0.0 ? h.h0 * exp(-V * h.hif) : h.h0);
}

__global__ void k_dep_hills(StepView v, HillBiasDev h, int w0, int nw)
{
    if (h.shared) {
        __shared__ int sN;
        if (threadIdx.x == 0) sN = h.count[0];
        __syncthreads();
        const int n0 = sN;
        for (int i = threadIdx.x; i < nw && n0 + i < h.cap; i += blockDim.x) deposit_hill(v, h, 0, w0 + i, n0 + i);
        __syncthreads();
        if (threadIdx.x == 0) h.count[0] = min(n0 + nw, h.cap > n0 ? h.cap : n0);
        return;
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nw) return;
    const int w = w0 + t;
    const int n = h.count[w];
    if (n >= h.cap) return;
    deposit_hill(v, h, w, w, n);
    h.count[w] = n + 1;
}

void launch_dep_hills(const StepView& v, const HillBiasDev& h, int w0, int nw, hipStream_t s)
{
    if (h.shared)
        hipLaunchKernelGGL(k_dep_hills, dim3(1), dim3(DEP_HILLS_SHARED_BLOCK), 0, s, v, h, w0, nw);
    else
        hipLaunchKernelGGL(k_dep_hills, dim3((nw + 63) / 64), dim3(64), 0, s, v, h, w0, nw);
}

// ---------------------------------------------------------------------------
// OPES deposit, CommonGluedKernels.cpp:2752-3003, one 1024-thread workgroup
// per kernel table.  The compression search is a parallel lexicographic
// (n2, index) arg-min over kernels with n2 < 1 -- identical to the
// reference's sequential strict-minimum scan because every kernel's n2 is
// formed with the same unfused FP64 operations.  sum_uprob is updated in
// O(N) on both append and merge: on a merge the old row and column of the
// merged kernel are subtracted and the new ones added, which equals the
// reference's O(N^2) recomputation up to FP64 rounding.
//
// Latency layout.  The table's first opes_lds_kernels<D>() kernels (centres,
// widths, log-weights; 4,000 at D = 2) are staged in LDS once per launch and
// kept in step with every merge and append, so the search, the merge and the
// sum_uprob update of a walker read no global memory; kernels past that are
// read from global memory where needed.  The scalar recursion (weights,
// n_eff, widths, height) and the merge arithmetic are formed redundantly by
// every thread from the same values, held in registers for the whole launch,
// so no broadcast is needed; thread 0 alone stores.  Two barriers per walker.
// The arithmetic and the summation order (kernel index ascending per thread,
// waves in order) are those of the one-lane scalar form, so results are
// unchanged.
// ---------------------------------------------------------------------------
constexpr int OPES_DEP_BLOCK = 512;
constexpr int OPES_DEP_LDS_BYTES = 160 * 1024 - 1024;  // the table stage; the rest holds the reductions

template <int D>
constexpr int opes_lds_kernels()
{
    return OPES_DEP_LDS_BYTES / (8 * (2 * D + 1));
}

// exp(-n2/2) - eps, scaled by the kernel height, for n2 below the cutoff
// (partial sums of n2 only grow, so stopping at the first one past the
// cutoff is the reference's test on the full sum).
template <int D>
__device__ __forceinline__ double opes_T(const double* ci, const double* ck, const double* sk, double hk,
                                         double cutoff2, double epsd)
{
#pragma clang fp contract(off)
    double n2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double dc = ci[d] - ck[d];
        n2 = n2 + dc * dc / (sk[d] * sk[d]);
        if (n2 >= cutoff2) return 0.0;
    }
    return hk * (exp(-0.5 * n2) - epsd);
}

// The same term with the kernel height given as its logarithm; the height
// exp(lk) is formed only inside the cutoff.
template <int D>
__device__ __forceinline__ double opes_Tl(const double* ci, const double* ck, const double* sk, double lk,
                                          double cutoff2, double epsd)
{
#pragma clang fp contract(off)
    double n2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double dc = ci[d] - ck[d];
        n2 = n2 + dc * dc / (sk[d] * sk[d]);
        if (n2 >= cutoff2) return 0.0;
    }
    return exp(lk) * (exp(-0.5 * n2) - epsd);
}

template <int D>
__global__ __launch_bounds__(OPES_DEP_BLOCK) void k_dep_opes(StepView v, OpesBiasDev o, int w0, int nw)
{
#pragma clang fp contract(off)
    constexpr int B = OPES_DEP_BLOCK, NWV = OPES_DEP_BLOCK / 64, LK = opes_lds_kernels<D>();
    __shared__ double tC[D * LK], tS[D * LK], tL[LK];  // staged table, [d][k]
    __shared__ double shN2[NWV], shSum[NWV];
    __shared__ int shIdx[NWV];
    const int tb = o.shared ? 0 : w0 + blockIdx.x;
    const int wFirst = o.shared ? w0 : w0 + blockIdx.x;
    const int wLast = o.shared ? w0 + nw : wFirst + 1;
    double* C = o.centers + (size_t)tb * D * o.cap;
    double* Sg = o.sigmas + (size_t)tb * D * o.cap;
    double* L = o.logw + (size_t)tb * o.cap;
    int* I = o.ints + tb * 4;
    double* Q = o.dbl + (size_t)tb * (4 + 2 * D);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int lk = min(o.cap, LK);

    // ---- the table scalars and the first walker's inputs, loaded together ----
    int i0 = I[0], i1 = I[1], i2 = I[2], i3 = I[3];
    double q0 = Q[0], q1 = Q[1], q2 = Q[2], q3 = Q[3], m2[D];
#pragma unroll
    for (int d = 0; d < D; d++) m2[d] = Q[4 + D + d];
    double Vb = v.benergy[(size_t)wFirst * v.B + o.eslot], cvx[D];
#pragma unroll
    for (int d = 0; d < D; d++) cvx[d] = v.cvVal[(size_t)wFirst * v.V + o.cv[d]];
    // ---- stage the kernels present into LDS ----
    const int nStage = min(i0, lk);
#pragma unroll 4
    for (int k = tid; k < nStage; k += B) {
        double c[D], sg[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            c[d] = C[(size_t)d * o.cap + k];
            sg[d] = Sg[(size_t)d * o.cap + k];
        }
        const double l = L[k];
#pragma unroll
        for (int d = 0; d < D; d++) {
            tC[d * LK + k] = c[d];
            tS[d * LK + k] = sg[d];
        }
        tL[k] = l;
    }
    const double epsd = exp(-0.5 * o.cutoff2);
    __syncthreads();

    for (int w = wFirst; w < wLast; w++) {
        if (w > wFirst) {
            Vb = v.benergy[(size_t)w * v.B + o.eslot];
#pragma unroll
            for (int d = 0; d < D; d++) cvx[d] = v.cvVal[(size_t)w * v.V + o.cv[d]];
        }
        // ---- weights, n_eff, widths, height (every thread, same values) ----
        if (o.adaptiveStride > 0 && i3 < o.adaptiveStride) continue;
        const int nB = i0;
        if (o.variant != 1 && o.adaptiveStride == 0) i3++;
        const int n = i3;
        double lw_raw, h_raw;
        if (o.variant == 2) {
            lw_raw = 0.0;
            h_raw = 1.0;
        } else {
            lw_raw = Vb * (1.0 / o.kT);
            h_raw = exp(lw_raw);
        }
        q2 += h_raw;
        q3 += h_raw * h_raw;
        i2++;
        const double neff = (1.0 + q2) * (1.0 + q2) / (1.0 + q3);
        const double size = o.variant == 2 ? (double)i2 : neff;
        double sC[D], sS[D], lw = lw_raw;
#pragma unroll
        for (int d = 0; d < D; d++) {
            double sg;
            if (o.variant == 1) {
                sg = o.sigma0[d];
            } else if (o.adaptiveStride > 0) {
                double factor = o.variant == 2 ? 1.0 : o.gamma;
                double var = n > 1 ? m2[d] / factor / (double)(n - 1) : 0.0;
                sg = var > 0.0 ? sqrt(var) : 1.0;
                if (sg < o.sigmaMin) sg = o.sigmaMin;
            } else {
                sg = o.sigma0[d] * pow(size * (D + 2) / 4.0, -1.0 / (D + 4.0));
                if (sg < o.sigmaMin) sg = o.sigmaMin;
            }
            sS[d] = sg;
            sC[d] = cvx[d];
        }
        if (o.variant != 1 && o.adaptiveStride == 0)
#pragma unroll
            for (int d = 0; d < D; d++) lw += log(o.sigma0[d] / sS[d]);
        const double h = exp(lw);
        q1 += h_raw;

        // ---- compression search: arg-min (n2, k) over n2 < 1 ----
        double best = 1.0;
        int bestK = -1;
        const int nS = min(nB, lk);
        for (int k = tid; k < nB; k += B) {  // k increases within a thread: first index wins ties
            double n2 = 0.0;
            if (k < nS) {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    double x = (sC[d] - tC[d * LK + k]) / tS[d * LK + k];
                    n2 = n2 + x * x;
                }
            } else {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    double x = (sC[d] - C[(size_t)d * o.cap + k]) / Sg[(size_t)d * o.cap + k];
                    n2 = n2 + x * x;
                }
            }
            if (n2 < best) {
                best = n2;
                bestK = k;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            double ob = __shfl_xor(best, off, 64);
            int ok = __shfl_xor(bestK, off, 64);
            if (ok >= 0 && (bestK < 0 || ob < best || (ob == best && ok < bestK))) {
                best = ob;
                bestK = ok;
            }
        }
        if (lane == 0) {
            shN2[wid] = best;
            shIdx[wid] = bestK;
        }
        __syncthreads();

        // ---- merge into the nearest kernel, or append (every thread forms
        // the same values; thread 0 stores) ----
        double b = 1.0;
        int mk = -1;
        for (int i = 0; i < NWV; i++) {
            const int ok = shIdx[i];
            if (ok >= 0 && (mk < 0 || shN2[i] < b || (shN2[i] == b && ok < mk))) {
                b = shN2[i];
                mk = ok;
            }
        }
        double oldC[D], oldS[D], newC[D], newS[D], newL = 0.0, oldH = 0.0, hNew = 0.0;
        int ka = -1;
        if (mk >= 0) {
            double oldL;
            if (mk < lk) {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    oldC[d] = tC[d * LK + mk];
                    oldS[d] = tS[d * LK + mk];
                }
                oldL = tL[mk];
            } else {
#pragma unroll
                for (int d = 0; d < D; d++) {
                    oldC[d] = C[(size_t)d * o.cap + mk];
                    oldS[d] = Sg[(size_t)d * o.cap + mk];
                }
                oldL = L[mk];
            }
            oldH = exp(oldL);
            const double lmax = oldL > lw ? oldL : lw, lmin = oldL < lw ? oldL : lw;
            newL = lmax + log1p(exp(lmin - lmax));
            const double ht = oldH + h;
            const double f1 = oldH / ht, f2 = h / ht;
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double c1 = oldC[d], s1 = oldS[d];
                newC[d] = f1 * c1 + f2 * sC[d];
                // the reference's f1(s1^2+c1^2) + f2(s2^2+c2^2) - cm^2
                // (CommonGluedKernels.cpp:2898-2905) in the algebraically
                // identical form without its cancellation at large |c|
                const double dcm = c1 - sC[d];
                double ss = f1 * s1 * s1 + f2 * sS[d] * sS[d] + f1 * f2 * dcm * dcm;
                if (ss < 0.0) ss = 0.0;
                double sm = sqrt(ss);
                if (sm < o.sigmaMin) sm = o.sigmaMin;
                newS[d] = sm;
            }
            hNew = exp(newL);
        } else if (i1 < o.cap) {
            ka = i1++;
            i0++;
        }
        // ---- sum_uprob update, O(N), against the table before this deposit ----
        double acc = 0.0;
        if (mk >= 0) {
            for (int i = tid; i < nB; i += B) {
                if (i == mk) {
                    // diagonal: row and column coincide
                    acc += opes_T<D>(newC, newC, newS, hNew, o.cutoff2, epsd) -
                           opes_T<D>(oldC, oldC, oldS, oldH, o.cutoff2, epsd);
                    continue;
                }
                double ci[D], si[D], li;
                if (i < nS) {
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        ci[d] = tC[d * LK + i];
                        si[d] = tS[d * LK + i];
                    }
                    li = tL[i];
                } else {
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        ci[d] = C[(size_t)d * o.cap + i];
                        si[d] = Sg[(size_t)d * o.cap + i];
                    }
                    li = L[i];
                }
                // row mk (j = mk, kernel i) and column mk (j = i, kernel mk)
                acc += opes_Tl<D>(newC, ci, si, li, o.cutoff2, epsd) - opes_Tl<D>(oldC, ci, si, li, o.cutoff2, epsd);
                acc += opes_T<D>(ci, newC, newS, hNew, o.cutoff2, epsd) -
                       opes_T<D>(ci, oldC, oldS, oldH, o.cutoff2, epsd);
            }
        } else if (ka >= 0) {
            for (int i = tid; i < nB; i += B) {
                double cj[D], sj[D], lj;
                if (i < nS) {
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        cj[d] = tC[d * LK + i];
                        sj[d] = tS[d * LK + i];
                    }
                    lj = tL[i];
                } else {
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        cj[d] = C[(size_t)d * o.cap + i];
                        sj[d] = Sg[(size_t)d * o.cap + i];
                    }
                    lj = L[i];
                }
                acc += opes_T<D>(cj, sC, sS, h, o.cutoff2, epsd);
                acc += opes_Tl<D>(sC, cj, sj, lj, o.cutoff2, epsd);
            }
        }
        acc = wsum(acc);
        if (lane == 0) shSum[wid] = acc;
        __syncthreads();
        // ---- store the changed kernel (global table and LDS stage) ----
        if (tid == 0 && (mk >= 0 || ka >= 0)) {
            const int kc = mk >= 0 ? mk : ka;
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double c = mk >= 0 ? newC[d] : sC[d], sg = mk >= 0 ? newS[d] : sS[d];
                C[(size_t)d * o.cap + kc] = c;
                Sg[(size_t)d * o.cap + kc] = sg;
                if (kc < lk) {
                    tC[d * LK + kc] = c;
                    tS[d * LK + kc] = sg;
                }
            }
            const double l = mk >= 0 ? newL : lw;
            L[kc] = l;
            if (kc < lk) tL[kc] = l;
        }
        if (mk >= 0 || ka >= 0) {
            double t = 0.0;
            for (int i = 0; i < NWV; i++) t += shSum[i];
            if (mk < 0) t += h * (1.0 - epsd);
            q0 += t;
        }
        __syncthreads();
    }
    if (tid == 0) {
        I[0] = i0;
        I[1] = i1;
        I[2] = i2;
        I[3] = i3;
        Q[0] = q0;
        Q[1] = q1;
        Q[2] = q2;
        Q[3] = q3;
    }
}

// Private tables: the same deposit spread over OPES_DEP_G workgroups per
// table in two launches.  k_dep_opes_find: every workgroup forms the scalar
// recursion (same values everywhere) and the arg-min over its kernels
// (kernel k belongs to workgroup (k / 256) mod G: lanes read consecutive
// kernels; the first two of a thread's kernels are loaded before anything
// else), and writes its best (n2, index) with that kernel's centre, width
// and log-weight; the last workgroup to finish (ticket) takes the
// lexicographic minimum of the G partials -- the same (n2, index) the
// one-workgroup search finds --, merges or appends, stores the table
// scalars, and records the change for the second launch.  k_dep_opes_sum:
// the O(N) sum_uprob change over the table as it was before the deposit, G
// partials added in workgroup order by the last workgroup.  Same terms as
// the one-workgroup form; only the grouping of the FP64 sum differs.
// Record per table (doubles): flag (0 none, 1 merge, 2 append), the changed
// kernel's index, nB, old height, new height, deposit height, then old
// centre, old width, new centre, new width, deposit centre, deposit width.
constexpr int OPES_SPLIT_BLOCK = 256, OPES_SPLIT_PF = 2;

template <int D>
__global__ __launch_bounds__(OPES_SPLIT_BLOCK) void k_dep_opes_find(StepView v, OpesBiasDev o, int w0)
{
#pragma clang fp contract(off)
    constexpr int G = OPES_DEP_G, NWV = OPES_SPLIT_BLOCK / 64, PF = OPES_SPLIT_PF, NV = 2 * D + 1;
    constexpr int STRIDE = G * OPES_SPLIT_BLOCK;
    __shared__ double shN2[NWV], shVal[NWV][NV];
    __shared__ int shIdx[NWV];
    const int tb = w0 + blockIdx.y, w = tb, g = blockIdx.x;
    double* C = o.centers + (size_t)tb * D * o.cap;
    double* Sg = o.sigmas + (size_t)tb * D * o.cap;
    double* L = o.logw + (size_t)tb * o.cap;
    int* I = o.ints + tb * 4;
    double* Q = o.dbl + (size_t)tb * (4 + 2 * D);
    double* part = o.depPart + (size_t)tb * G * OPES_PART;
    double* rec = o.depRec + (size_t)tb * OPES_REC;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int k0 = g * OPES_SPLIT_BLOCK + tid;

    // ---- every load the search needs first: this thread's first kernels
    // (clamped into the capacity), the table scalars, the walker's inputs ----
    double pc[PF][D], ps[PF][D], pl[PF];
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const int k = min(k0 + j * STRIDE, o.cap - 1);
#pragma unroll
        for (int d = 0; d < D; d++) {
            pc[j][d] = C[(size_t)d * o.cap + k];
            ps[j][d] = Sg[(size_t)d * o.cap + k];
        }
        pl[j] = L[k];
    }
    int i0 = I[0], i1 = I[1], i2 = I[2], i3 = I[3];
    double q1 = Q[1], q2 = Q[2], q3 = Q[3], m2[D];
#pragma unroll
    for (int d = 0; d < D; d++) m2[d] = Q[4 + D + d];
    const double Vb = v.benergy[(size_t)w * v.B + o.eslot];
    double cvx[D];
#pragma unroll
    for (int d = 0; d < D; d++) cvx[d] = v.cvVal[(size_t)w * v.V + o.cv[d]];
    if (o.adaptiveStride > 0 && i3 < o.adaptiveStride) {
        if (g == 0 && tid == 0) rec[0] = 0.0;
        return;
    }
    const int nB = i0;
    if (o.variant != 1 && o.adaptiveStride == 0) i3++;
    const int n = i3;
    double lw_raw, h_raw;
    if (o.variant == 2) {
        lw_raw = 0.0;
        h_raw = 1.0;
    } else {
        lw_raw = Vb * (1.0 / o.kT);
        h_raw = exp(lw_raw);
    }
    q2 += h_raw;
    q3 += h_raw * h_raw;
    i2++;
    const double neff = (1.0 + q2) * (1.0 + q2) / (1.0 + q3);
    const double size = o.variant == 2 ? (double)i2 : neff;
    double sC[D], sS[D], lw = lw_raw;
#pragma unroll
    for (int d = 0; d < D; d++) {
        double sg;
        if (o.variant == 1) {
            sg = o.sigma0[d];
        } else if (o.adaptiveStride > 0) {
            double factor = o.variant == 2 ? 1.0 : o.gamma;
            double var = n > 1 ? m2[d] / factor / (double)(n - 1) : 0.0;
            sg = var > 0.0 ? sqrt(var) : 1.0;
            if (sg < o.sigmaMin) sg = o.sigmaMin;
        } else {
            sg = o.sigma0[d] * pow(size * (D + 2) / 4.0, -1.0 / (D + 4.0));
            if (sg < o.sigmaMin) sg = o.sigmaMin;
        }
        sS[d] = sg;
        sC[d] = cvx[d];
    }
    if (o.variant != 1 && o.adaptiveStride == 0)
#pragma unroll
        for (int d = 0; d < D; d++) lw += log(o.sigma0[d] / sS[d]);
    q1 += h_raw;

    // ---- this workgroup's share of the arg-min (n2, k) over n2 < 1 ----
    double best = 1.0, bv[NV];
    int bestK = -1;
#pragma unroll
    for (int e = 0; e < NV; e++) bv[e] = 0.0;
#pragma unroll
    for (int j = 0; j < PF; j++) {
        double n2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            double x = (sC[d] - pc[j][d]) / ps[j][d];
            n2 = n2 + x * x;
        }
        if (k0 + j * STRIDE < nB && n2 < best) {  // k increases within a thread: first index wins ties
            best = n2;
            bestK = k0 + j * STRIDE;
#pragma unroll
            for (int d = 0; d < D; d++) {
                bv[d] = pc[j][d];
                bv[D + d] = ps[j][d];
            }
            bv[2 * D] = pl[j];
        }
    }
    for (int k = k0 + PF * STRIDE; k < nB; k += STRIDE) {
        double c[D], sg[D], n2 = 0.0;
#pragma unroll
        for (int d = 0; d < D; d++) {
            c[d] = C[(size_t)d * o.cap + k];
            sg[d] = Sg[(size_t)d * o.cap + k];
        }
        const double l = L[k];
#pragma unroll
        for (int d = 0; d < D; d++) {
            double x = (sC[d] - c[d]) / sg[d];
            n2 = n2 + x * x;
        }
        if (n2 < best) {
            best = n2;
            bestK = k;
#pragma unroll
            for (int d = 0; d < D; d++) {
                bv[d] = c[d];
                bv[D + d] = sg[d];
            }
            bv[2 * D] = l;
        }
    }
    const int myK = bestK;
    for (int off = 32; off > 0; off >>= 1) {
        double ob = __shfl_xor(best, off, 64);
        int ok = __shfl_xor(bestK, off, 64);
        if (ok >= 0 && (bestK < 0 || ob < best || (ob == best && ok < bestK))) {
            best = ob;
            bestK = ok;
        }
    }
    if (lane == 0) {
        shN2[wid] = best;
        shIdx[wid] = bestK;
    }
    if (bestK >= 0 && myK == bestK)  // the lane holding the wave's best kernel
#pragma unroll
        for (int e = 0; e < NV; e++) shVal[wid][e] = bv[e];
    __syncthreads();
    if (tid == 0) {
        double b = 1.0;
        int bk = -1, bw = 0;
        for (int i = 0; i < NWV; i++) {
            const int ok = shIdx[i];
            if (ok >= 0 && (bk < 0 || shN2[i] < b || (shN2[i] == b && ok < bk))) {
                b = shN2[i];
                bk = ok;
                bw = i;
            }
        }
        double* pp = part + g * OPES_PART;
        pp[0] = b;
        pp[1] = (double)bk;
#pragma unroll
        for (int e = 0; e < NV; e++) pp[2 + e] = shVal[bw][e];
    }
    if (!last_block(&o.depCount[2 * tb], G)) return;
    if (tid >= 64) return;

    // ---- last workgroup, wave 0: lane i reads partial i; the minimum ----
    double b = 1.0, mv[NV];
    int mk = -1;
#pragma unroll
    for (int e = 0; e < NV; e++) mv[e] = 0.0;
    if (lane < G) {
        const double* pp = part + lane * OPES_PART;
        b = __hip_atomic_load(pp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mk = (int)__hip_atomic_load(pp + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int e = 0; e < NV; e++) mv[e] = __hip_atomic_load(pp + 2 + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (mk < 0) b = 1.0;
    }
    const int ownK = mk;
    for (int off = 32; off > 0; off >>= 1) {
        double ob = __shfl_xor(b, off, 64);
        int ok = __shfl_xor(mk, off, 64);
        if (ok >= 0 && (mk < 0 || ob < b || (ob == b && ok < mk))) {
            b = ob;
            mk = ok;
        }
    }
    // the winning lane's kernel values, to every lane
    const unsigned long long own = __ballot(mk >= 0 && ownK == mk);
    const int src = own ? __ffsll(own) - 1 : 0;
#pragma unroll
    for (int e = 0; e < NV; e++) mv[e] = __shfl(mv[e], src, 64);
    if (lane != 0) return;
    const double h = exp(lw);
    double flag = 0.0, kc = -1.0, oldH = 0.0, hNew = 0.0;
    if (mk >= 0) {
        double oldC[D], oldS[D], newC[D], newS[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            oldC[d] = mv[d];
            oldS[d] = mv[D + d];
        }
        const double oldL = mv[2 * D];
        oldH = exp(oldL);
        const double lmax = oldL > lw ? oldL : lw, lmin = oldL < lw ? oldL : lw;
        const double newL = lmax + log1p(exp(lmin - lmax));
        const double ht = oldH + h;
        const double f1 = oldH / ht, f2 = h / ht;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const double c1 = oldC[d], s1 = oldS[d];
            newC[d] = f1 * c1 + f2 * sC[d];
            // the reference's f1(s1^2+c1^2) + f2(s2^2+c2^2) - cm^2
            // (CommonGluedKernels.cpp:2898-2905) in the algebraically
            // identical form without its cancellation at large |c|
            const double dcm = c1 - sC[d];
            double ss = f1 * s1 * s1 + f2 * sS[d] * sS[d] + f1 * f2 * dcm * dcm;
            if (ss < 0.0) ss = 0.0;
            double sm = sqrt(ss);
            if (sm < o.sigmaMin) sm = o.sigmaMin;
            newS[d] = sm;
        }
        hNew = exp(newL);
#pragma unroll
        for (int d = 0; d < D; d++) {
            C[(size_t)d * o.cap + mk] = newC[d];
            Sg[(size_t)d * o.cap + mk] = newS[d];
            rec[6 + d] = oldC[d];
            rec[6 + D + d] = oldS[d];
            rec[6 + 2 * D + d] = newC[d];
            rec[6 + 3 * D + d] = newS[d];
        }
        L[mk] = newL;
        flag = 1.0;
        kc = mk;
    } else if (i1 < o.cap) {
        const int ka = i1++;
        i0++;
#pragma unroll
        for (int d = 0; d < D; d++) {
            C[(size_t)d * o.cap + ka] = sC[d];
            Sg[(size_t)d * o.cap + ka] = sS[d];
        }
        L[ka] = lw;
        flag = 2.0;
        kc = ka;
    }
#pragma unroll
    for (int d = 0; d < D; d++) {
        rec[6 + 4 * D + d] = sC[d];
        rec[6 + 5 * D + d] = sS[d];
    }
    rec[0] = flag;
    rec[1] = kc;
    rec[2] = nB;
    rec[3] = oldH;
    rec[4] = hNew;
    rec[5] = h;
    I[0] = i0;
    I[1] = i1;
    I[2] = i2;
    I[3] = i3;
    Q[1] = q1;
    Q[2] = q2;
    Q[3] = q3;
    o.depCount[2 * tb] = 0u;
}

template <int D>
__global__ __launch_bounds__(OPES_SPLIT_BLOCK) void k_dep_opes_sum(OpesBiasDev o, int w0)
{
#pragma clang fp contract(off)
    constexpr int G = OPES_DEP_G, NWV = OPES_SPLIT_BLOCK / 64, PF = OPES_SPLIT_PF;
    constexpr int STRIDE = G * OPES_SPLIT_BLOCK;
    __shared__ double shSum[NWV];
    const int tb = w0 + blockIdx.y, g = blockIdx.x;
    const double* C = o.centers + (size_t)tb * D * o.cap;
    const double* Sg = o.sigmas + (size_t)tb * D * o.cap;
    const double* L = o.logw + (size_t)tb * o.cap;
    double* part = o.depPart + (size_t)tb * G * OPES_PART;
    const double* rec = o.depRec + (size_t)tb * OPES_REC;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int k0 = g * OPES_SPLIT_BLOCK + tid;
    // this thread's first kernels (clamped into the capacity) with the record
    double pc[PF][D], ps[PF][D], pl[PF];
#pragma unroll
    for (int j = 0; j < PF; j++) {
        const int k = min(k0 + j * STRIDE, o.cap - 1);
#pragma unroll
        for (int d = 0; d < D; d++) {
            pc[j][d] = C[(size_t)d * o.cap + k];
            ps[j][d] = Sg[(size_t)d * o.cap + k];
        }
        pl[j] = L[k];
    }
    const double flag = rec[0];
    const int mk = (int)rec[1], nB = (int)rec[2];
    const double oldH = rec[3], hNew = rec[4], h = rec[5];
    double oldC[D], oldS[D], newC[D], newS[D], sC[D], sS[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        oldC[d] = rec[6 + d];
        oldS[d] = rec[6 + D + d];
        newC[d] = rec[6 + 2 * D + d];
        newS[d] = rec[6 + 3 * D + d];
        sC[d] = rec[6 + 4 * D + d];
        sS[d] = rec[6 + 5 * D + d];
    }
    if (flag == 0.0) return;
    const double epsd = exp(-0.5 * o.cutoff2);
    // the change's terms against kernel i (before the deposit)
    auto terms = [&](int i, const double* ci, const double* si, double li) -> double {
        if (flag == 1.0) {
            if (i == mk)  // diagonal: row and column coincide
                return opes_T<D>(newC, newC, newS, hNew, o.cutoff2, epsd) -
                       opes_T<D>(oldC, oldC, oldS, oldH, o.cutoff2, epsd);
            // row mk (j = mk, kernel i) and column mk (j = i, kernel mk)
            double a = opes_Tl<D>(newC, ci, si, li, o.cutoff2, epsd) - opes_Tl<D>(oldC, ci, si, li, o.cutoff2, epsd);
            return a + (opes_T<D>(ci, newC, newS, hNew, o.cutoff2, epsd) -
                        opes_T<D>(ci, oldC, oldS, oldH, o.cutoff2, epsd));
        }
        double a = opes_T<D>(ci, sC, sS, h, o.cutoff2, epsd);
        return a + opes_Tl<D>(sC, ci, si, li, o.cutoff2, epsd);
    };
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (k0 + j * STRIDE < nB) acc += terms(k0 + j * STRIDE, pc[j], ps[j], pl[j]);
    for (int i = k0 + PF * STRIDE; i < nB; i += STRIDE) {
        double ci[D], si[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            ci[d] = C[(size_t)d * o.cap + i];
            si[d] = Sg[(size_t)d * o.cap + i];
        }
        acc += terms(i, ci, si, L[i]);
    }
    acc = wsum(acc);
    if (lane == 0) shSum[wid] = acc;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int i = 0; i < NWV; i++) t += shSum[i];
        part[g * OPES_PART] = t;
    }
    if (!last_block(&o.depCount[2 * tb + 1], G)) return;
    if (tid != 0) return;
    double t = 0.0;
    for (int i = 0; i < G; i++)
        t += __hip_atomic_load(part + i * OPES_PART, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (flag == 2.0) t += h * (1.0 - epsd);
    double* Q = o.dbl + (size_t)tb * (4 + 2 * D);
    Q[0] += t;
    o.depCount[2 * tb + 1] = 0u;
}

// Every-step Welford accumulator (adaptive sigma), :3009-3027.
__global__ void k_opes_welford(StepView v, OpesBiasDev o, int w0, int nw)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (o.shared ? t != 0 : t >= nw) return;
    const int wFirst = o.shared ? w0 : w0 + t;
    const int wLast = o.shared ? w0 + nw : wFirst + 1;
    const int tb = o.shared ? 0 : wFirst;
    int* I = o.ints + tb * 4;
    double* Q = o.dbl + (size_t)tb * (4 + 2 * o.D);
    for (int w = wFirst; w < wLast; w++) {
        const double* cv = v.cvVal + (size_t)w * v.V;
        I[3]++;
        const int n = I[3];
        for (int d = 0; d < o.D; d++) {
            double x = cv[o.cv[d]];
            double delta = x - Q[4 + d];
            Q[4 + d] += delta / n;
            Q[4 + o.D + d] += delta * (x - Q[4 + d]);
        }
    }
}

void launch_dep_opes(const StepView& v, const OpesBiasDev& o, int w0, int nw, hipStream_t s)
{
#define GLUED_DEP_OPES(DD)                                                                                          \
    if (o.shared) {                                                                                                  \
        hipLaunchKernelGGL(k_dep_opes<DD>, dim3(1), dim3(OPES_DEP_BLOCK), 0, s, v, o, w0, nw);                        \
    } else {                                                                                                         \
        hipLaunchKernelGGL(k_dep_opes_find<DD>, dim3(OPES_DEP_G, nw), dim3(OPES_SPLIT_BLOCK), 0, s, v, o, w0);       \
        hipLaunchKernelGGL(k_dep_opes_sum<DD>, dim3(OPES_DEP_G, nw), dim3(OPES_SPLIT_BLOCK), 0, s, o, w0);           \
    }
    switch (o.D) {
    case 1: GLUED_DEP_OPES(1) break;
    case 2: GLUED_DEP_OPES(2) break;
    case 3: GLUED_DEP_OPES(3) break;
    case 4: GLUED_DEP_OPES(4) break;
    case 5: GLUED_DEP_OPES(5) break;
    case 6: GLUED_DEP_OPES(6) break;
    case 7: GLUED_DEP_OPES(7) break;
    default: GLUED_DEP_OPES(8) break;
    }
#undef GLUED_DEP_OPES
}

void launch_opes_welford(const StepView& v, const OpesBiasDev& o, int w0, int nw, hipStream_t s)
{
    const int n = o.shared ? 1 : nw;
    hipLaunchKernelGGL(k_opes_welford, dim3((n + 63) / 64), dim3(64), 0, s, v, o, w0, nw);
}

// OPES expanded logZ update (:3094-3125) and multithermal dF update
// (:3174-3203); one thread per (item, walker).
// One wave per (item, walker).  OPES-expanded: lane 0 (the log-sum-exp over
// the states in state order).  Multithermal: the states' recursions are
// independent, lanes over states (every lane forms the walker's scalars
// from the values read at entry; lane 0 writes them back last).
__global__ void k_dep_small(StepView v, SmallBiasState st, const SmallBiasItem* __restrict__ items, int n, int w0,
                            int onlyItem)
{
    const int w = w0 + blockIdx.y;
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= n || (onlyItem >= 0 && i != onlyItem)) return;
    const SmallBiasItem it = items[i];
    const double* cv = v.cvVal + (size_t)w * v.V;
    double* sc = st.scal + (size_t)w * st.scalStride + 4 * i;
    if (it.type == BIAS_OPES_EXPANDED) {
        if (lane != 0) return;
        const double invKT = it.p[0];
        const double* lw = st.stateConst + it.stateOff;
        const double* ci = lw + it.nState;
        double mx = -1e300;
        for (int l = 0; l < it.nState; l++) mx = fmax(mx, lw[l] - cv[(int)ci[l]] * invKT);
        double se = 0.0;
        for (int l = 0; l < it.nState; l++) se += exp(lw[l] - cv[(int)ci[l]] * invKT - mx);
        const double logQ = mx + log(se);
        const int k = (int)sc[1];
        if (k == 0) {
            sc[0] = logQ;
        } else {
            double A = sc[0] + log((double)k), B = logQ, m = A > B ? A : B;
            sc[0] = m + log(exp(A - m) + exp(B - m)) - log((double)(k + 1));
        }
        sc[1] = k + 1;
    } else if (it.type == BIAS_OPES_MULTITHERMAL) {
        const double kT0 = it.p[0];
        const double* db = st.stateConst + it.stateOff;
        double* dF = st.stateVar + (size_t)w * st.stateVarStride + it.stateOff;
        const double U = cv[it.cv[0]];
        const double V = v.benergy[(size_t)w * v.B + it.eslot];
        const double r = sc[2];
        const double cnt = sc[3] + 1.0, logc = log(cnt - 1.0);
        const double arg = (V - r) / kT0 - logc;
        const double inc = arg > 0.0 ? kT0 * (arg + log1p(exp(-arg))) : kT0 * log1p(exp(arg));
        for (int l = lane; l < it.nState; l += 64) {
            double dl = -db[l] * U + (V - r + dF[l]) / kT0 - logc;
            double dec = dl > 0.0 ? kT0 * (dl + log1p(exp(-dl))) : kT0 * log1p(exp(dl));
            dF[l] += inc - dec;
        }
        if (lane == 0) {
            sc[3] = cnt;
            sc[2] = r + inc + kT0 * log1p(-1.0 / cnt);
        }
    }
}

void launch_dep_small(const StepView& v, const SmallBiasState& st, const SmallBiasItem* items, int n, int item,
                      int w0, int nw, hipStream_t s)
{
    hipLaunchKernelGGL(k_dep_small, dim3(n, nw), dim3(64), 0, s, v, st, items, n, w0, item);
}

}  // namespace glued

namespace glued {

// ---------------------------------------------------------------------------
// Adjoint of the hill-list bias with respect to the hill heights and widths,
// for the autograd path (no counterpart in the reference):
//   dL/dh_k     = sum_w g_w exp(-E_wk) [E_wk <= 8]
//   dL/dsigma_d = sum_w g_w sum_k h_k exp(-E_wk) [E_wk <= 8] dc_wkd^2 / sigma_d^3
// with the forward's exact arithmetic (FP64 exponent from the same host
// reciprocals, FP32 exp).  Shared table: one thread per hill, walkers' CV
// values and g_w staged through LDS in tiles of 256.  Private tables: one
// y-row of blocks per walker.
// ---------------------------------------------------------------------------
constexpr int ADJ_BLOCK = 256;

template <int D>
__global__ __launch_bounds__(ADJ_BLOCK) void k_hill_adjoint(StepView v, HillBiasDev h, const double* __restrict__ gV,
                                                            double* __restrict__ gH, double* __restrict__ sigPart,
                                                            int W)
{
#pragma clang fp contract(off)
    __shared__ double sS[ADJ_BLOCK * D];
    __shared__ double sG[ADJ_BLOCK];
    __shared__ double sh[4 * D];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int tb = h.shared ? 0 : blockIdx.y;
    const int n = h.count[tb];
    const float* C = h.centers + (size_t)tb * D * h.cap;
    double c[D], hk = 0.0;
    const bool ok = k < n;
#pragma unroll
    for (int d = 0; d < D; d++) c[d] = ok ? (double)C[(size_t)d * h.cap + k] : 0.0;
    if (ok) hk = (double)h.heights[(size_t)tb * h.cap + k];
    double gh = 0.0, gs[D];
#pragma unroll
    for (int d = 0; d < D; d++) gs[d] = 0.0;
    const int wBegin = h.shared ? 0 : blockIdx.y;
    const int wEnd = h.shared ? W : blockIdx.y + 1;
    for (int w0 = wBegin; w0 < wEnd; w0 += ADJ_BLOCK) {
        const int m = min(ADJ_BLOCK, wEnd - w0);
        if (threadIdx.x < m) {
            const double* cv = v.cvVal + (size_t)(w0 + threadIdx.x) * v.V;
#pragma unroll
            for (int d = 0; d < D; d++) sS[threadIdx.x * D + d] = cv[h.cv[d]];
            sG[threadIdx.x] = gV[w0 + threadIdx.x];
        }
        __syncthreads();
        if (ok)
            for (int j = 0; j < m; j++) {
                double e = 0.0, dc[D];
#pragma unroll
                for (int d = 0; d < D; d++) {
                    double x = sS[j * D + d] - c[d];
                    if (h.period[d] > 0.0) x = x - rint(x * h.invPeriod[d]) * h.period[d];
                    dc[d] = x;
                    e = e + x * x * h.i2s2[d];
                }
                if (!(e <= 8.0)) continue;
                const double t = sG[j] * hill_exp(e);
                gh += t;
#pragma unroll
                for (int d = 0; d < D; d++) gs[d] += t * hk * dc[d] * dc[d];
            }
        __syncthreads();
    }
    if (ok) gH[(size_t)tb * h.cap + k] = gh;
    // per-block partial of the sigma gradient (scaled by 1/sigma^3 on the host side of the reduction)
    wave_sum_n<D>(gs);
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < D; d++) sh[d * 4 + (threadIdx.x >> 6)] = gs[d];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int d = 0; d < D; d++)
            sigPart[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * D + d] =
                sh[d * 4] + sh[d * 4 + 1] + sh[d * 4 + 2] + sh[d * 4 + 3];
}

// Sum the per-block sigma partials in block order: gSigma[d] = sum / sigma_d^3.
__global__ void k_hill_adjoint_sigma(HillBiasDev h, const double* sigPart, int nPart, double* gSigma)
{
    const int d = threadIdx.x;
    if (d >= h.D) return;
    double s = 0.0;
    for (int i = 0; i < nPart; i++) s += sigPart[(size_t)i * h.D + d];
    gSigma[d] = s / (h.sigma[d] * h.sigma[d] * h.sigma[d]);
}

int hill_adjoint_blocks(int cap) { return (cap + ADJ_BLOCK - 1) / ADJ_BLOCK; }

void launch_hill_adjoint(const StepView& v, const HillBiasDev& h, const double* gV, double* gH, double* sigPart,
                         double* gSigma, int W, hipStream_t s)
{
    dim3 g(hill_adjoint_blocks(h.cap), h.shared ? 1 : W), b(ADJ_BLOCK);
    switch (h.D) {
        case 1: hipLaunchKernelGGL(k_hill_adjoint<1>, g, b, 0, s, v, h, gV, gH, sigPart, W); break;
        case 2: hipLaunchKernelGGL(k_hill_adjoint<2>, g, b, 0, s, v, h, gV, gH, sigPart, W); break;
        case 3: hipLaunchKernelGGL(k_hill_adjoint<3>, g, b, 0, s, v, h, gV, gH, sigPart, W); break;
        default: hipLaunchKernelGGL(k_hill_adjoint<4>, g, b, 0, s, v, h, gV, gH, sigPart, W); break;
    }
    hipLaunchKernelGGL(k_hill_adjoint_sigma, dim3(1), dim3(64), 0, s, h, sigPart, (int)(g.x * g.y), gSigma);
}

// FP64 device values to FP32 storage (hill heights given by the autograd
// wrapper), round to nearest.
__global__ void k_f64_to_f32(const double* __restrict__ src, float* __restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

void launch_f64_to_f32(const double* src, float* dst, int n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_f64_to_f32, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, n);
}

// ---------------------------------------------------------------------------
// Adjoint of an OPES bias with respect to its kernel parameters (an
// extension: the reference has no parameter derivatives).  For the eval of
// CommonGluedKernels.cpp:2642-2723, V = A log(max(pz + eps, 1e-30)),
// A = invGF kT, pz = P n / U, P = sum_k h_k (G_k - Gfloor) over the kernels
// it includes (n2 < cutoff2, G_k > Gfloor, finite h_k):
//   F_w            = g_w dV_w/dP = g_w A n / (U (pz + eps))   (pz / P = n / U),
//                    with pz + eps = exp(V_w / A) from the walker's last V,
//   dL/dlogw_k    += F_w h_k (G_k - Gfloor),
//   dL/dc_kd      += F_w h_k G_k (s_d - c_kd) / sigma_kd^2,
//   dL/dsigma_kd  += F_w h_k G_k (s_d - c_kd)^2 / sigma_kd^3,
//   dL/dU         += -g_w A pz / (U (pz + eps))   (sum_uprob, the stored normaliser),
// summed over the walkers that read the table; zero where the eval's
// log-argument floor or the no-kernel branch applies.  One thread per kernel,
// the walkers' CV values and factors staged in LDS, walkers in index order.
// Outputs: gC, gS [T][cap][D] kernel-major, gL [T][cap], gU [T].
constexpr int OPES_ADJ_BLOCK = 256;

template <int D>
__global__ __launch_bounds__(OPES_ADJ_BLOCK) void k_opes_adjoint(StepView v, OpesBiasDev o,
                                                                 const double* __restrict__ gV, double* gC,
                                                                 double* gS, double* gL, double* gU, int W)
{
#pragma clang fp contract(off)
    __shared__ double sS[OPES_ADJ_BLOCK][D];
    __shared__ double sF[OPES_ADJ_BLOCK], sFU[OPES_ADJ_BLOCK];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int tb = o.shared ? 0 : blockIdx.y;
    const int nk = o.ints[tb * 4 + 0];
    const double su = o.dbl[(size_t)tb * (4 + 2 * D) + 0];
    const double A = o.invGF * o.kT;
    const double kfloor = exp(-0.5 * o.cutoff2);
    const bool ok = k < nk;
    const int kc = min(k, max(nk - 1, 0));
    double c[D], sg[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        c[d] = o.centers[((size_t)tb * D + d) * o.cap + kc];
        sg[d] = o.sigmas[((size_t)tb * D + d) * o.cap + kc];
    }
    const double lw = o.logw[(size_t)tb * o.cap + kc];
    double gc[D], gs[D], gl = 0.0, gu = 0.0;
#pragma unroll
    for (int d = 0; d < D; d++) gc[d] = gs[d] = 0.0;
    const int wBegin = o.shared ? 0 : blockIdx.y;
    const int wEnd = o.shared ? W : blockIdx.y + 1;
    for (int w0 = wBegin; w0 < wEnd; w0 += OPES_ADJ_BLOCK) {
        const int m = min(OPES_ADJ_BLOCK, wEnd - w0);
        __syncthreads();
        if ((int)threadIdx.x < m) {
            const int w = w0 + threadIdx.x;
            const double* cv = v.cvVal + (size_t)w * v.V;
#pragma unroll
            for (int d = 0; d < D; d++) sS[threadIdx.x][d] = cv[o.cv[d]];
            double F = 0.0, FU = 0.0;
            if (nk > 0 && su > 0.0) {
                const double varg = exp(v.benergy[(size_t)w * v.B + o.eslot] / A);  // pz + eps of the last eval
                if (varg > 1e-30 * (1.0 + 1e-9)) {
                    F = gV[w] * A * (double)nk / (su * varg);
                    FU = -gV[w] * A * (varg - o.eps) / (su * varg);
                }
            }
            sF[threadIdx.x] = F;
            sFU[threadIdx.x] = FU;
        }
        __syncthreads();
        if (blockIdx.x == 0 && threadIdx.x == 0)
            for (int j = 0; j < m; j++) gu += sFU[j];
        if (!ok) continue;
        for (int j = 0; j < m; j++) {
            const double F = sF[j];
            if (F == 0.0) continue;
            double n2 = 0.0, dc[D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                dc[d] = sS[j][d] - c[d];
                n2 = n2 + dc[d] * dc[d] / (sg[d] * sg[d]);
            }
            if (!(n2 < o.cutoff2)) continue;
            const double G = exp(-0.5 * n2), kern = G - kfloor;
            if (kern <= 0.0) continue;
            const double hk = exp(lw);
            if (!(hk > 0.0 && hk < 1e308)) continue;
            const double t = F * hk;
            gl += t * kern;
#pragma unroll
            for (int d = 0; d < D; d++) {
                const double q = t * G * dc[d] / (sg[d] * sg[d]);
                gc[d] += q;
                gs[d] += q * dc[d] / sg[d];
            }
        }
    }
    if (k < o.cap) {
#pragma unroll
        for (int d = 0; d < D; d++) {
            gC[((size_t)tb * o.cap + k) * D + d] = ok ? gc[d] : 0.0;
            gS[((size_t)tb * o.cap + k) * D + d] = ok ? gs[d] : 0.0;
        }
        gL[(size_t)tb * o.cap + k] = ok ? gl : 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) gU[tb] = gu;
}

void launch_opes_adjoint(const StepView& v, const OpesBiasDev& o, const double* gV, double* gC, double* gS,
                         double* gL, double* gU, int W, hipStream_t s)
{
    const dim3 g((o.cap + OPES_ADJ_BLOCK - 1) / OPES_ADJ_BLOCK, o.shared ? 1 : W), b(OPES_ADJ_BLOCK);
    switch (o.D) {
        case 1: hipLaunchKernelGGL(k_opes_adjoint<1>, g, b, 0, s, v, o, gV, gC, gS, gL, gU, W); break;
        case 2: hipLaunchKernelGGL(k_opes_adjoint<2>, g, b, 0, s, v, o, gV, gC, gS, gL, gU, W); break;
        case 3: hipLaunchKernelGGL(k_opes_adjoint<3>, g, b, 0, s, v, o, gV, gC, gS, gL, gU, W); break;
        default: hipLaunchKernelGGL(k_opes_adjoint<4>, g, b, 0, s, v, o, gV, gC, gS, gL, gU, W); break;
    }
}

}  // namespace glued
