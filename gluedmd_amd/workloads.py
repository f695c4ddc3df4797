"""Seeded synthetic workloads for the five benchmark configurations.

Random numbers: SplitMix64 (state += 0x9E3779B97F4A7C15; xor-shift-multiply
finaliser), U[0,1) = (x >> 11) * 2^-53, N(0, s) by Box-Muller (cos branch,
u1 taken as 1 - U so it is never 0).  Seeds are 0x474C554544000000 + config
number.  Positions are emitted as FP32; every consumer (GPU engine, oracle)
sees the identical FP32 values.

T = 300 K (kT = 2.494338785 kJ/mol) wherever a configuration leaves T open.
Builder choices that the configurations leave open are written next to each
generator.
"""
import math

import numpy as np

from .glued import Force, multithermal_temperature_ladder

SEED_BASE = 0x474C554544000000
KT300 = 8.314462618e-3 * 300.0
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * _GOLD
            self.state = self.state + np.uint64(n) * _GOLD
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.next_u64(int(n)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n, sigma=1.0):
        n = int(n)
        u1 = 1.0 - self.uniform(n)
        u2 = self.uniform(n)
        return sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)

    def integers(self, n, lo, hi):
        """Uniform integers in [lo, hi] inclusive."""
        return (lo + np.floor(self.uniform(n) * (hi - lo + 1))).astype(np.int64)


def fp32_values(a):
    """FP64 array of FP32-representable values (hill centres and heights are
    stored in FP32)."""
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


def rng(config):
    return SplitMix64(SEED_BASE + int(config))


# ---------------------------------------------------------------------------
# Config 1: plain method, small (the reference's CPU-runnable case)
# ---------------------------------------------------------------------------
def config1(n_frames=1000, n_prior=1000):
    """22 atoms U[0,3)^3 in a 3.0 nm periodic cube; CVs d(0,21), phi(4,6,8,14),
    psi(6,8,14,16); well-tempered MetaD (hill list) on (phi, psi): sigma 0.35
    rad, h0 1.2 kJ/mol, gamma 6, T 300 K, period 2 pi; 1,000 prior hills with
    centres U[-pi,pi)^2 and heights U[0.1,1.2]; frames = base + N(0, 0.01 nm)."""
    r = rng(1)
    base = r.uniform(22 * 3, 0.0, 3.0).reshape(22, 3)
    frames = np.stack([base + r.normal(22 * 3, 0.01).reshape(22, 3) for _ in range(n_frames)]).astype(np.float32)
    hc = fp32_values(r.uniform(2 * n_prior, -math.pi, math.pi).reshape(n_prior, 2))
    hh = fp32_values(r.uniform(n_prior, 0.1, 1.2))
    f = Force(pbc=True, temperature=300.0)
    f.add_distance([0, 21])
    phi = f.add_dihedral([4, 6, 8, 14])
    psi = f.add_dihedral([6, 8, 14, 16])
    f.add_metad_hills([phi, psi], sigma=0.35, height=1.2, pace=1, gamma=6.0, period=2 * math.pi,
                      max_hills=n_prior + n_frames + 8)
    return dict(force=f, natoms=22, box=3.0, frames=frames, hill_centers=hc, hill_heights=hh, bias=0)


# ---------------------------------------------------------------------------
# Config 2: coordination + gyration, OPES-METAD, one walker
# ---------------------------------------------------------------------------
def config2_system():
    """50,000 atoms U[0,L)^3 at 100 atoms/nm^3 (L = 500^(1/3) nm)."""
    r = rng(2)
    N = 50000
    L = (N / 100.0) ** (1.0 / 3.0)
    pos = r.uniform(3 * N, 0.0, L).reshape(N, 3).astype(np.float32)
    f = Force(pbc=True, temperature=300.0)
    cn = f.add_coordination(list(range(0, 2000)), list(range(2000, 22000)), r0=0.4, n=6, m=12, cutoff=1.2)
    rg = f.add_gyration(list(range(0, 5000)))
    gamma = 40.0 / KT300  # BARRIER = 40 kJ/mol
    f.add_opes([cn, rg], sigma=[0.05, 0.02], gamma=gamma, pace=1, temperature=300.0, max_kernels=65536)
    return dict(force=f, natoms=N, box=L, positions=pos, gamma=gamma, rng=r)


def opes_prior_kernels(r, center, sigma0, gamma, n=5000, spread=20.0):
    """Builder choice: n kernels with centres center + sigma0 * N(0, spread),
    sigma = sigma0, log-weights U[-gamma, 0]."""
    D = len(center)
    c = np.asarray(center)[None, :] + np.asarray(sigma0)[None, :] * r.normal(n * D, spread).reshape(n, D)
    s = np.tile(np.asarray(sigma0, dtype=np.float64), (n, 1))
    lw = r.uniform(n, -gamma, 0.0)
    return c, s, lw


def opes_metad_cutoff2(gamma):
    """Squared kernel cutoff of OPES-METAD with bias factor gamma:
    2 gamma / invGF with invGF = (gamma - 1) / gamma (CommonBiasKernels.cpp:553-588)."""
    return 2.0 * gamma * gamma / (gamma - 1.0)


def opes_state_blob(centers, sigmas, logw, sum_uprob, n_metad=0, grids=()):
    """A 'GPUS' v1 bias-state blob holding one OPES bias
    (CommonSerialization.cpp:20-178 layout)."""
    import struct
    n, D = centers.shape
    b = bytearray(b"GPUS") + struct.pack("<i", 1) + struct.pack("<i", 1)
    b += struct.pack("<i", n) + struct.pack("<d", sum_uprob) + struct.pack("<i", 0)
    b += np.zeros(2 * D).tobytes()
    b += np.ascontiguousarray(centers, dtype="<f8").tobytes()
    b += np.ascontiguousarray(sigmas, dtype="<f8").tobytes()
    b += np.ascontiguousarray(logw, dtype="<f8").tobytes()
    b += struct.pack("<i", 0)  # ABMD
    b += struct.pack("<i", n_metad)
    for nd, g in grids:
        b += struct.pack("<i", nd) + np.ascontiguousarray(g, dtype="<f8").tobytes()
    return bytes(b)


def sum_uprob_np(centers, sigmas, logw, cutoff2, block=512):
    """O(N^2) sum_uprob = sum_j sum_k h_k (exp(-n2/2) - eps)[n2 < cutoff2] with
    n2 using kernel k's sigma (CommonGluedKernels.cpp:2980-3002), in numpy."""
    eps = math.exp(-0.5 * cutoff2)
    h = np.exp(logw)
    tot = 0.0
    for j0 in range(0, len(h), block):
        dc = centers[j0:j0 + block, None, :] - centers[None, :, :]
        n2 = np.sum(dc * dc / (sigmas[None, :, :] ** 2), axis=2)
        t = h[None, :] * (np.exp(-0.5 * n2) - eps)
        tot += float(np.sum(np.where(n2 < cutoff2, t, 0.0)))
    return tot


# ---------------------------------------------------------------------------
# Config 3: differentiable path-CV -> MetaD pipeline
# ---------------------------------------------------------------------------
def config3_system(batch=256, M=300, N=64, n_hills=20000):
    """Reference of 300 atoms U[0,3)^3; 64 frames = random walk from it with
    N(0, 0.03 nm) steps; 256 configurations = reference + N(0, 0.05 nm);
    lambda 50 nm^-2.  Hill centres are drawn once the (s, z) range of the
    configurations is known (config3_hills)."""
    r = rng(3)
    ref = r.uniform(3 * M, 0.0, 3.0).reshape(M, 3)
    frames = [ref.copy()]
    for _ in range(N - 1):
        frames.append(frames[-1] + r.normal(3 * M, 0.03).reshape(M, 3))
    frames = np.stack(frames).astype(np.float32)
    confs = np.stack([ref + r.normal(3 * M, 0.05).reshape(M, 3) for _ in range(batch)]).astype(np.float32)
    return dict(ref=ref, frames=frames, confs=confs, lam=50.0, M=M, N=N, rng=r, n_hills=n_hills)


def config3_hills(r, s_range, z_range, n_hills=20000):
    """Builder choice: hill centres uniform over the (s, z) bounding box of the
    configurations, heights U[0.1,1.2], sigma_s = 0.5, sigma_z = 0.1 z-range."""
    c = fp32_values(np.stack([r.uniform(n_hills, *s_range), r.uniform(n_hills, *z_range)], axis=1))
    h = fp32_values(r.uniform(n_hills, 0.1, 1.2))
    sig = np.array([0.5, 0.1 * max(z_range[1] - z_range[0], 1e-3)])
    return c, h, sig


# ---------------------------------------------------------------------------
# Config 4: batched multi-walker, uneven sizes
# ---------------------------------------------------------------------------
def random_rotation(r):
    q = r.normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def config4(n_walkers=64, n_hills=20000, extra_hills=4096):
    """64 walkers, N_w ~ U{2000..40000}, cubic box at 100 atoms/nm^3.
    CVs: Kabsch RMSD of atoms [0,500) to a seeded 500-atom reference and
    torsions (500..503), (504..507).  Bias: WT MetaD on (RMSD, phi1) with a
    shared 200x200 grid (RMSD [0,2] nm non-periodic, phi periodic) plus 20,000
    private hills per walker.
    Builder choices: reference atoms U[0,2)^3; walker atoms [0,500) = random
    rotation of the reference + N(0, 0.1 nm) noise, translated into the box;
    other atoms U[0,L)^3; sigma (0.05 nm, 0.35 rad), h0 1.2, gamma 6, pace 1;
    hill centres RMSD U[0,0.6], phi U[-pi,pi), heights U[0.1,1.2]; the grid is
    pre-filled with 400 Gaussians at centres drawn the same way."""
    r = rng(4)
    sizes = r.integers(n_walkers, 2000, 40000)
    ref = r.uniform(500 * 3, 0.0, 2.0).reshape(500, 3)
    refc = ref - ref.mean(axis=0)
    pos = []
    for w in range(n_walkers):
        N = int(sizes[w])
        L = (N / 100.0) ** (1.0 / 3.0)
        p = r.uniform(3 * N, 0.0, L).reshape(N, 3)
        R = random_rotation(r)
        p[:500] = refc @ R.T + r.normal(1500, 0.1).reshape(500, 3) + L / 2
        pos.append(p.astype(np.float32))
    hills = []
    for w in range(n_walkers):
        c = np.stack([r.uniform(n_hills, 0.0, 0.6), r.uniform(n_hills, -math.pi, math.pi)], axis=1)
        hills.append((fp32_values(c), fp32_values(r.uniform(n_hills, 0.1, 1.2))))
    f = Force(pbc=True, temperature=300.0)
    rmsd = f.add_rmsd(list(range(500)), ref.tolist(), fit=True)
    phi = f.add_dihedral([500, 501, 502, 503])
    f.add_dihedral([504, 505, 506, 507])
    grid_bias = f.add_metad([rmsd, phi], sigma=[0.05, 0.35], height=1.2, pace=1, grid_min=[0.0, -math.pi],
                            grid_max=[2.0, math.pi], bins=[199, 200], periodic=[False, True], gamma=6.0)
    f.share_bias(grid_bias)
    hill_bias = f.add_metad_hills([rmsd, phi], sigma=[0.05, 0.35], height=1.2, pace=1, gamma=6.0,
                                  period=[0.0, 2 * math.pi], max_hills=n_hills + extra_hills)
    prior = np.stack([r.uniform(400, 0.0, 0.6), r.uniform(400, -math.pi, math.pi)], axis=1)
    return dict(force=f, sizes=[int(s) for s in sizes], positions=pos, ref=ref, hills=hills,
                boxes=[(int(s) / 100.0) ** (1.0 / 3.0) for s in sizes], grid_bias=grid_bias, hill_bias=hill_bias,
                grid_prior=prior)


def grid_prefill(force, grid_bias, centers, height=1.2):
    """Grid values of `centers` deposited with the grid rule (flat height),
    computed in numpy (used to give the shared grid non-trivial content)."""
    t, cvs, params, ip = force.biases[grid_bias]
    D = len(cvs)
    sigma = np.array(params[1:1 + D])
    mn = np.array(params[3 + D:3 + 2 * D])
    mx = np.array(params[3 + 2 * D:3 + 3 * D])
    bins = np.array(ip[1:1 + D])
    per = np.array(ip[1 + D:1 + 2 * D])
    npts = bins + np.where(per != 0, 0, 1)
    axes = [mn[d] + np.arange(npts[d]) * (mx[d] - mn[d]) / bins[d] for d in range(D)]
    mesh = np.meshgrid(*axes, indexing="ij")  # mesh[d][i0, i1]
    grid = np.zeros(tuple(npts))
    for c in centers:
        e = np.zeros_like(grid)
        for d in range(D):
            dc = mesh[d] - c[d]
            if per[d]:
                P = npts[d] * (mx[d] - mn[d]) / bins[d]
                dc = dc - np.round(dc / P) * P
            e += dc * dc / (2 * sigma[d] ** 2)
        grid += np.where(e <= 8.0, height * np.exp(-e), 0.0)
    # column-major flat index: strides[0] = 1
    return grid.reshape(-1, order="F")


# ---------------------------------------------------------------------------
# Config 5: large mixed workload
# ---------------------------------------------------------------------------
def config5(n_walkers=16, N=100000, n_cm=4096, n_drmsd=1500, path_frames=256, path_atoms=1000, n_temps=32):
    """100,000 atoms at 100 atoms/nm^3 (L = 10 nm); contact map over 4,096
    seeded pairs (r0 0.5, n 6, m 12, w 1); DRMSD over 1,500 pairs of a seeded
    reference with d_ref < 0.8 nm; path CV 256 frames x 1,000 atoms (frames =
    random walk, N(0, 0.03 nm) steps, from walker 0's atoms; lambda 100 nm^-2;
    the other walkers' path atoms = walker 0's + N(0, 0.05 nm)); multithermal
    on a seeded energy CV, 32 temperatures 300-500 K (geometric); 16 walkers.
    Builder choices: the DRMSD reference is the first walker's positions;
    U ~ N(-1.5e6, 1e3) kJ/mol and F ~ N(0, 500) kJ/mol/nm per walker."""
    r = rng(5)
    L = (N / 100.0) ** (1.0 / 3.0)
    pos = [r.uniform(3 * N, 0.0, L).reshape(N, 3).astype(np.float32) for _ in range(n_walkers)]
    cm_pairs = r.integers(2 * n_cm, 0, N - 1).reshape(n_cm, 2)
    ref = pos[0].astype(np.float64)
    # candidate pairs (a, b) are consecutive draws of one stream, kept while
    # d_ref < 0.8 nm; drawn in blocks, with the stream rewound to just after
    # the last candidate used
    pairs, d0 = [], []
    while len(pairs) < n_drmsd:
        start = r.state
        cand = r.integers(2 * 65536, 0, N - 1).reshape(-1, 2)
        d = ref[cand[:, 1]] - ref[cand[:, 0]]
        d -= L * np.round(d / L)
        dist = np.sqrt(np.sum(d * d, axis=1))
        keep = np.nonzero((cand[:, 0] != cand[:, 1]) & (dist < 0.8))[0][:n_drmsd - len(pairs)]
        pairs += [(int(cand[i, 0]), int(cand[i, 1])) for i in keep]
        d0 += [float(np.linalg.norm(d[i])) for i in keep]
        if len(pairs) == n_drmsd:
            used = 2 * (int(keep[-1]) + 1)
            with np.errstate(over="ignore"):
                r.state = start + np.uint64(used) * _GOLD
    path_idx = list(range(path_atoms))
    base = pos[0][:path_atoms].astype(np.float64)
    frames = [base.copy()]
    for _ in range(path_frames - 1):
        frames.append(frames[-1] + r.normal(3 * path_atoms, 0.03).reshape(path_atoms, 3))
    for w in range(1, n_walkers):  # every walker's path atoms sit near the path
        pos[w][:path_atoms] = (base + r.normal(3 * path_atoms, 0.05).reshape(path_atoms, 3)).astype(np.float32)
    U = r.normal(n_walkers, 1e3) - 1.5e6
    F = [r.normal(3 * N, 500.0).reshape(N, 3).astype(np.float32) for _ in range(n_walkers)]
    f = Force(pbc=True, temperature=300.0)
    f.add_contact_map(cm_pairs.reshape(-1).tolist(), r0=0.5, n=6, m=12, w=1.0)
    f.add_drmsd([x for p in pairs for x in p], d0)
    f.add_path(path_idx, [fr.reshape(-1).tolist() for fr in frames], 100.0)
    f.add_multithermal(300.0, 500.0, n_temps, temp0=300.0, pace=1)
    return dict(force=f, natoms=N, box=L, positions=pos, U=U, F=F,
                ladder=multithermal_temperature_ladder(300.0, 500.0, n_temps))
