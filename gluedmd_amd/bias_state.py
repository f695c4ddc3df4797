"""The 'GPUS' version-1 bias-state blob on the host: parse, pack, merge.

Layout (little-endian), in the order of CommonSerialization.cpp:20-178 (the
writer) and :180-410 (the reader):

    'GPUS', int32 version = 1          (absent in legacy blobs)
    int32 nOpes;  per bias: int32 nk, f64 sum_uprob (the field the reference
                  calls logZ), int32 nSamples, f64[D] mean, f64[D] M2,
                  f64[nk*D] centres, f64[nk*D] widths, f64[nk] log-weights
    int32 nAbmd;  per bias: f64[D] rhoMin
    int32 nMetaD; per bias: int32 numDeposited, f64[G] grid
    int32 nPb;    per bias: int32 nSub; per sub-grid: int32 numDeposited, f64[G_i]
    int32 nExternal, int32 nLinear, int32 nWall      (counts only)
    int32 nOpesExpanded; per bias: f64 logZ, int32 numUpdates
    int32 nExtLag; per bias: f64[D] s, f64[D] p
    int32 nEds;    per bias: f64[D] lambda, mean, ssd, accum; int32[D] count
    int32 nMaxEnt; per bias: f64[D] lambda
    int32 nMultithermal; per bias: int32 N, f64[N] deltaF, f64 rct, f64 counter
    'HILL' extension of this library: int32 nLists; per list: int32 n,
           int32 D, f64[n*D] centres, f64[n] heights

Every section after the OPES one is optional at the end of a blob (the
reference reader stops where the bytes stop), so a parsed blob remembers how
many sections it held and packs back to the same bytes.

Dimensions come from the force definition (`getNumBiases` / `getBiasInfo`,
type codes of GluedForce.h): the blob stores none.
Grid sizes follow the kernels' parameter layouts: MetaD int parameters are
[pace, bins_0..bins_{D-1}, periodic_0..periodic_{D-1}]
(CommonBiasKernels.cpp:206-233), PBMetaD ones [pace, bins_0, periodic_0,
bins_1, ...] (:298-330); a grid axis has bins + 1 points, bins if periodic.

BiasStateMerger follows MultiGPUManager.py:132-420: MetaD and PBMetaD grids
add (incrementally against per-group baselines after the first sync), OPES
kernel lists are concatenated with the walkers' Welford statistics combined
by Chan's parallel formula, OPES-expanded log Z is averaged in linear space
weighted by the update counts, and every other section is taken from the
first blob.  Hill lists (the extension) are concatenated the same way.
"""
import math
import struct

import numpy as np

MAGIC = b"GPUS"
HILL_MAGIC = b"HILL"
VERSION = 1

# bias type codes, GluedForce.h:72-94
_METAD, _PBMETAD, _OPES, _ABMD = 3, 4, 5, 7
_OPES_EXPANDED, _EXT_LAG, _MAXENT, _EDS, _MULTITHERMAL = 11, 12, 13, 14, 15
_HILLS = 101


def _bias(force, i):
    """(type, cv indices, params, int params) of bias i: GluedForce::getBiasInfo
    (GluedForce.h:128), or the getBiasParameters spelling of the Python proxy."""
    get = getattr(force, "getBiasInfo", None) or force.getBiasParameters
    return get(i)

# the optional sections after OPES, in blob order
SECTIONS = ("abmd", "metad", "pbmetad", "external", "linear_wall", "opes_expanded", "ext_lag", "eds", "maxent",
            "multithermal", "hills")


class BlobError(ValueError):
    """A bias-state blob that does not match its force definition."""


class ForceLayout:
    """Per-section dimensions of a force's biases (what the blob omits)."""

    def __init__(self, force):
        self.opes, self.abmd, self.metad, self.pbmetad = [], [], [], []
        self.ext_lag, self.eds, self.maxent, self.hills = [], [], [], []
        self.n_opes_expanded = self.n_multithermal = 0
        for i in range(force.getNumBiases()):
            t, cvs, params, ip = _bias(force, i)
            D = len(cvs)
            if t == _OPES:
                self.opes.append(D)
            elif t == _ABMD:
                self.abmd.append(D)
            elif t == _METAD:
                G = 1
                for d in range(D):
                    G *= ip[1 + d] + (0 if ip[1 + D + d] else 1)
                self.metad.append(G)
            elif t == _PBMETAD:
                self.pbmetad.append([ip[1 + 2 * d] + (0 if ip[2 + 2 * d] else 1) for d in range(D)])
            elif t == _EXT_LAG:
                self.ext_lag.append(D)
            elif t == _EDS:
                self.eds.append(D)
            elif t == _MAXENT:
                self.maxent.append(D)
            elif t == _OPES_EXPANDED:
                self.n_opes_expanded += 1
            elif t == _MULTITHERMAL:
                self.n_multithermal += 1
            elif t == _HILLS:
                self.hills.append(D)


class OpesTable:
    """One OPES bias's state: kernel table, sum_uprob and Welford statistics."""

    def __init__(self, sum_uprob, n_samples, mean, m2, centers, sigmas, logw):
        self.sum_uprob = float(sum_uprob)
        self.n_samples = int(n_samples)
        self.mean = np.asarray(mean, dtype=np.float64).reshape(-1)
        self.m2 = np.asarray(m2, dtype=np.float64).reshape(-1)
        D = len(self.mean)
        self.centers = np.asarray(centers, dtype=np.float64).reshape(-1, D)
        self.sigmas = np.asarray(sigmas, dtype=np.float64).reshape(-1, D)
        self.logw = np.asarray(logw, dtype=np.float64).reshape(-1)

    @property
    def nk(self):
        return len(self.logw)


class BiasState:
    """A parsed blob.  `n_sections` counts the optional sections present
    after the OPES one (SECTIONS order); `header` is False for a legacy blob."""

    def __init__(self):
        self.header = True
        self.opes = []            # OpesTable
        self.abmd = []            # f64[D]
        self.metad = []           # (numDeposited, f64[G])
        self.pbmetad = []         # [(numDeposited, f64[G_i]) per sub-grid]
        self.n_external = self.n_linear = self.n_wall = 0
        self.opes_expanded = []   # (logZ, numUpdates)
        self.ext_lag = []         # (s f64[D], p f64[D])
        self.eds = []             # (lambda, mean, ssd, accum f64[D], count i32[D])
        self.maxent = []          # f64[D]
        self.multithermal = []    # (deltaF f64[N], rct, counter)
        self.hills = []           # (centres f64[n, D], heights f64[n])
        self.n_sections = len(SECTIONS)


class _Reader:
    def __init__(self, blob):
        self.b = memoryview(bytes(blob))
        self.p = 0

    def left(self):
        return len(self.b) - self.p

    def take(self, n):
        if n < 0 or self.p + n > len(self.b):
            raise BlobError(f"bias-state blob truncated: {n} bytes wanted at offset {self.p}, {self.left()} left")
        out = self.b[self.p:self.p + n]
        self.p += n
        return out

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def f64s(self, n):
        return np.frombuffer(self.take(8 * n), dtype="<f8").astype(np.float64)

    def i32s(self, n):
        return np.frombuffer(self.take(4 * n), dtype="<i4").astype(np.int32)


def _count(r, what, expected):
    n = r.i32()
    if expected is not None and n != expected:
        raise BlobError(f"bias-state {what} count: the blob holds {n}, the force defines {expected}")
    return n


def parse(blob, force):
    """Parse a bias-state blob against the force that wrote it."""
    lay = ForceLayout(force)
    r = _Reader(blob)
    st = BiasState()
    if r.left() >= 8 and bytes(r.b[:4]) == MAGIC:
        r.take(4)
        v = r.i32()
        if v != VERSION:
            raise BlobError(f"bias-state blob version {v}; this library reads version {VERSION}")
    else:
        st.header = False
    for D in (lay.opes[i] for i in range(_count(r, "OPES", len(lay.opes)))):
        nk = r.i32()
        if nk < 0:
            raise BlobError(f"negative OPES kernel count {nk}")
        su, ns = r.f64(), r.i32()
        mean, m2 = r.f64s(D), r.f64s(D)
        st.opes.append(OpesTable(su, ns, mean, m2, r.f64s(nk * D), r.f64s(nk * D), r.f64s(nk)))
    done = 0
    for sec in SECTIONS:
        if r.left() == 0:
            break
        if sec == "abmd":
            st.abmd = [r.f64s(D) for D in lay.abmd[:_count(r, "ABMD", len(lay.abmd))]]
        elif sec == "metad":
            st.metad = [(r.i32(), r.f64s(G)) for G in lay.metad[:_count(r, "MetaD", len(lay.metad))]]
        elif sec == "pbmetad":
            for sizes in lay.pbmetad[:_count(r, "PBMetaD", len(lay.pbmetad))]:
                nsub = _count(r, "PBMetaD sub-grid", len(sizes))
                st.pbmetad.append([(r.i32(), r.f64s(G)) for G in sizes[:nsub]])
        elif sec == "external":
            st.n_external = r.i32()
        elif sec == "linear_wall":
            st.n_linear = r.i32()
            st.n_wall = r.i32()
        elif sec == "opes_expanded":
            n = _count(r, "OPES_EXPANDED", lay.n_opes_expanded)
            st.opes_expanded = [(r.f64(), r.i32()) for _ in range(n)]
        elif sec == "ext_lag":
            st.ext_lag = [(r.f64s(D), r.f64s(D)) for D in lay.ext_lag[:_count(r, "EXT_LAGRANGIAN", len(lay.ext_lag))]]
        elif sec == "eds":
            st.eds = [(r.f64s(D), r.f64s(D), r.f64s(D), r.f64s(D), r.i32s(D))
                      for D in lay.eds[:_count(r, "EDS", len(lay.eds))]]
        elif sec == "maxent":
            st.maxent = [r.f64s(D) for D in lay.maxent[:_count(r, "MAXENT", len(lay.maxent))]]
        elif sec == "multithermal":
            for _ in range(_count(r, "OPES_MULTITHERMAL", lay.n_multithermal)):
                N = r.i32()
                st.multithermal.append((r.f64s(N), r.f64(), r.f64()))
        elif sec == "hills":
            if bytes(r.take(4)) != HILL_MAGIC:
                raise BlobError("unknown section after the multithermal one")
            for D in lay.hills[:_count(r, "hill-list", len(lay.hills))]:
                n, d2 = r.i32(), r.i32()
                if d2 != D:
                    raise BlobError(f"hill list of {d2} CVs; the force defines {D}")
                st.hills.append((r.f64s(n * D).reshape(n, D), r.f64s(n)))
        done += 1
    if r.left():
        raise BlobError(f"{r.left()} bytes after the last section")
    st.n_sections = done
    return st


def pack(st):
    """The blob of a BiasState (inverse of parse, byte for byte)."""
    out = bytearray()

    def i32(x):
        out.extend(struct.pack("<i", int(x)))

    def f64(x):
        out.extend(struct.pack("<d", float(x)))

    def arr(a, dt="<f8"):
        out.extend(np.ascontiguousarray(a, dtype=dt).tobytes())

    if st.header:
        out.extend(MAGIC)
        i32(VERSION)
    i32(len(st.opes))
    for o in st.opes:
        i32(o.nk)
        f64(o.sum_uprob)
        i32(o.n_samples)
        arr(o.mean)
        arr(o.m2)
        arr(o.centers)
        arr(o.sigmas)
        arr(o.logw)
    for sec in SECTIONS[:st.n_sections]:
        if sec == "abmd":
            i32(len(st.abmd))
            for a in st.abmd:
                arr(a)
        elif sec == "metad":
            i32(len(st.metad))
            for nd, g in st.metad:
                i32(nd)
                arr(g)
        elif sec == "pbmetad":
            i32(len(st.pbmetad))
            for subs in st.pbmetad:
                i32(len(subs))
                for nd, g in subs:
                    i32(nd)
                    arr(g)
        elif sec == "external":
            i32(st.n_external)
        elif sec == "linear_wall":
            i32(st.n_linear)
            i32(st.n_wall)
        elif sec == "opes_expanded":
            i32(len(st.opes_expanded))
            for lz, nu in st.opes_expanded:
                f64(lz)
                i32(nu)
        elif sec == "ext_lag":
            i32(len(st.ext_lag))
            for s, p in st.ext_lag:
                arr(s)
                arr(p)
        elif sec == "eds":
            i32(len(st.eds))
            for lam, mean, ssd, acc, cnt in st.eds:
                for a in (lam, mean, ssd, acc):
                    arr(a)
                arr(cnt, "<i4")
        elif sec == "maxent":
            i32(len(st.maxent))
            for a in st.maxent:
                arr(a)
        elif sec == "multithermal":
            i32(len(st.multithermal))
            for dF, rct, cnt in st.multithermal:
                i32(len(dF))
                arr(dF)
                f64(rct)
                f64(cnt)
        elif sec == "hills":
            out.extend(HILL_MAGIC)
            i32(len(st.hills))
            for c, h in st.hills:
                i32(len(h))
                i32(c.shape[1] if c.ndim == 2 else 0)
                arr(c)
                arr(h)
    return bytes(out)


# ---------------------------------------------------------------------------
# merging (MultiGPUManager.py:132-420)
# ---------------------------------------------------------------------------
def _chan(tables):
    """Welford statistics of several walkers combined pairwise (Chan et al.):
    delta = mean_b - mean_a; mean += delta n_b / n; M2 += M2_b + delta^2 n_a n_b / n.
    Walkers with no samples are skipped."""
    n, mean, m2 = 0, None, None
    for t in tables:
        if t.n_samples <= 0:
            continue
        if n == 0:
            n, mean, m2 = t.n_samples, t.mean.copy(), t.m2.copy()
            continue
        tot = n + t.n_samples
        delta = t.mean - mean
        mean = mean + delta * (t.n_samples / tot)
        m2 = m2 + t.m2 + delta * delta * (n * t.n_samples / tot)
        n = tot
    if n == 0:
        D = len(tables[0].mean)
        return 0, np.zeros(D), np.zeros(D)
    return n, mean, m2


def _logsumexp(x):
    m = float(np.max(x))
    return m + math.log(float(np.sum(np.exp(x - m))))


def opes_cutoff2(variant, gamma):
    """Squared kernel cutoff of an OPES bias (CommonBiasKernels.cpp:553-588)."""
    if variant in (1, 2):
        return 2.0 * gamma
    inv_gf = (gamma - 1.0) / gamma
    return 2.0 * gamma / inv_gf


def sum_uprob(centers, sigmas, logw, cutoff2, block=256):
    """sum_j sum_k exp(logw_k) (exp(-n2_jk / 2) - exp(-cutoff2 / 2)) over
    n2_jk < cutoff2, n2_jk = |(c_j - c_k) / sigma_k|^2: the table's
    normalisation sum as the deposit forms it (CommonGluedKernels.cpp:2980-3002)."""
    eps = math.exp(-0.5 * cutoff2)
    h = np.exp(logw)
    tot = 0.0
    for j0 in range(0, len(h), block):
        dc = centers[j0:j0 + block, None, :] - centers[None, :, :]
        n2 = np.sum(dc * dc / (sigmas[None, :, :] ** 2), axis=2)
        tot += float(np.sum(np.where(n2 < cutoff2, h[None, :] * (np.exp(-0.5 * n2) - eps), 0.0)))
    return tot


class BiasStateMerger:
    """Merge the bias states of several walker groups (one blob per group)."""

    @staticmethod
    def merge_additive(blobs, force, opes_sum_uprob="reference"):
        """First-sync merge: grids summed in full, kernel and hill lists
        concatenated (MultiGPUManager.py:154-181)."""
        return BiasStateMerger.merge_additive_incremental(blobs, force, None, opes_sum_uprob)[0]

    @staticmethod
    def merge_additive_incremental(blobs, force, baselines=None, opes_sum_uprob="reference"):
        """(merged blob, new baselines).  With one baseline per group (the
        parsed state each group last received), grid values and deposit
        counts add only each group's change since then:
        new = shared + sum_g (G_g - base_g), shared = baselines[0]
        (MultiGPUManager.py:184-300); hill lists gain each group's hills
        past its baseline's count.  Otherwise every blob adds in full.

        opes_sum_uprob: "reference" stores log sum_k exp(logw_k) of the
        union in the OPES table's sum_uprob field, as the reference merger
        does (MultiGPUManager.py:376-383, which takes the field for a log
        normaliser); "recompute" stores the union's pairwise sum_uprob, the
        value the deposit kernels keep there."""
        if opes_sum_uprob not in ("reference", "recompute"):
            raise ValueError("opes_sum_uprob must be 'reference' or 'recompute'")
        if not blobs:
            raise ValueError("no bias-state blobs to merge")
        states = [parse(b, force) for b in blobs]
        if len(states) == 1:
            return bytes(blobs[0]), [states[0]]
        inc = baselines is not None and len(baselines) == len(states)
        out = states[0]
        res = BiasState()
        res.__dict__.update(out.__dict__)

        def add(cur, base_all, shared):
            if not inc:
                return sum(cur[1:], cur[0].copy())
            tot = shared.copy()
            for c, b in zip(cur, base_all):
                tot = tot + (c - b)
            return tot

        res.metad = []
        for i in range(len(out.metad)):
            grids = [s.metad[i][1] for s in states]
            nds = [s.metad[i][0] for s in states]
            if inc:
                g = add(grids, [b.metad[i][1] for b in baselines], baselines[0].metad[i][1])
                nd = baselines[0].metad[i][0] + sum(n - b.metad[i][0] for n, b in zip(nds, baselines))
            else:
                g, nd = add(grids, None, None), sum(nds)
            res.metad.append((int(nd), g))
        res.pbmetad = []
        for i in range(len(out.pbmetad)):
            subs = []
            for j in range(len(out.pbmetad[i])):
                grids = [s.pbmetad[i][j][1] for s in states]
                nds = [s.pbmetad[i][j][0] for s in states]
                if inc:
                    g = add(grids, [b.pbmetad[i][j][1] for b in baselines], baselines[0].pbmetad[i][j][1])
                    nd = baselines[0].pbmetad[i][j][0] + sum(n - b.pbmetad[i][j][0] for n, b in zip(nds, baselines))
                else:
                    g, nd = add(grids, None, None), sum(nds)
                subs.append((int(nd), g))
            res.pbmetad.append(subs)
        res.opes = []
        opes_params = [_bias(force, i) for i in range(force.getNumBiases())]
        opes_params = [(p, ip) for t, _, p, ip in opes_params if t == _OPES]
        for i in range(len(out.opes)):
            tabs = [s.opes[i] for s in states]
            n, mean, m2 = _chan(tabs)
            logw = np.concatenate([t.logw for t in tabs])
            cen = np.concatenate([t.centers for t in tabs])
            sig = np.concatenate([t.sigmas for t in tabs])
            if not len(logw):
                su = tabs[0].sum_uprob
            elif opes_sum_uprob == "reference":
                su = _logsumexp(logw)
            else:
                p, ip = opes_params[i]
                su = sum_uprob(cen, sig, logw, opes_cutoff2(ip[0] if ip else 0, p[1]))
            res.opes.append(OpesTable(su, n, mean, m2, cen, sig, logw))
        res.opes_expanded = []
        for i in range(len(out.opes_expanded)):
            ent = [s.opes_expanded[i] for s in states]
            nu = sum(int(e[1]) for e in ent)
            if nu > 0:
                m = max(e[0] for e in ent)
                acc = sum(e[1] * math.exp(e[0] - m) for e in ent if e[1] > 0)
                res.opes_expanded.append((m + math.log(acc / nu), nu))
            else:
                res.opes_expanded.append((sum(e[0] for e in ent) / len(ent), 0))
        res.hills = []
        for i in range(len(out.hills)):
            if inc:
                c = [baselines[0].hills[i][0]] + [s.hills[i][0][len(b.hills[i][1]):] for s, b in zip(states, baselines)]
                h = [baselines[0].hills[i][1]] + [s.hills[i][1][len(b.hills[i][1]):] for s, b in zip(states, baselines)]
            else:
                c = [s.hills[i][0] for s in states]
                h = [s.hills[i][1] for s in states]
            res.hills.append((np.concatenate(c), np.concatenate(h)))
        merged = pack(res)
        again = parse(merged, force)
        return merged, [again for _ in states]
