"""The 'GPUS' v1 bias-state blob on the host (gluedmd_amd.bias_state): every
section of CommonSerialization.cpp:20-178 parsed field by field from bytes
assembled here with struct, byte-exact repacking (also of blobs that end early
and of headerless legacy blobs), and the BiasStateMerger rules of
MultiGPUManager.py:132-420, including the reference tests' cases
(tests/test_multi_gpu.py:470-575) and closed forms for the Welford and log Z
combinations."""
import math
import struct

import numpy as np
import pytest

from gluedmd_amd import bias_state as bs
from gluedmd_amd.glued import Force


class _Layout:
    """A force definition reduced to what the blob needs: (type, cvs, params, ints)."""

    def __init__(self, biases):
        self.biases = biases

    def getNumBiases(self):
        return len(self.biases)

    def getBiasInfo(self, i):
        return self.biases[i]


def i32(x):
    return struct.pack("<i", x)


def f64(*xs):
    return b"".join(struct.pack("<d", x) for x in xs)


def metad_blob(grid, nd=0):
    """The reference test's one-MetaD blob (tests/test_multi_gpu.py:470-487)."""
    b = b"GPUS" + i32(1) + i32(0) + i32(0) + i32(1) + i32(nd) + f64(*grid) + i32(0)
    return b + i32(0) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0)


META1 = _Layout([(3, [0], [], [10, 5, 0])])  # MetaD, 1 CV, 5 bins, not periodic: 6 points


def test_reference_merge_two_metad_blobs():
    a, b = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0], [0.5] * 6
    merged = bs.BiasStateMerger.merge_additive([metad_blob(a, 10), metad_blob(b, 5)], META1)
    st = bs.parse(merged, META1)
    assert st.metad[0][0] == 15
    np.testing.assert_allclose(st.metad[0][1], np.array(a) + np.array(b), rtol=1e-12)


def test_reference_single_and_three_blobs():
    blob = metad_blob([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], 3)
    assert bs.BiasStateMerger.merge_additive([blob], META1) == blob
    blobs = [metad_blob([float(k)] * 6, 10) for k in (1, 2, 3)]
    st = bs.parse(bs.BiasStateMerger.merge_additive(blobs, META1), META1)
    np.testing.assert_allclose(st.metad[0][1], np.full(6, 6.0), rtol=1e-12)
    assert st.metad[0][0] == 30


def test_reference_repack_roundtrip():
    blob = metad_blob([0.1 * i for i in range(6)], 7)
    assert bs.pack(bs.parse(blob, META1)) == blob


def _full_force():
    """One bias of every kind whose state the blob carries."""
    return _Layout([
        (5, [0, 1], [2.5, 10.0, 0.1, 0.1, 0.01], [0, 1, 100]),     # OPES, D = 2
        (7, [2], [1.0], []),                                       # ABMD, D = 1
        (3, [0, 1], [], [10, 4, 3, 1, 0]),                         # MetaD: bins (4, 3), periodic (1, 0): 4 x 4
        (4, [0, 2], [], [10, 5, 0, 6, 1]),                         # PBMetaD: sub-grids of 6 and 6 points
        (10, [0], [1.0], []),                                      # linear
        (8, [0], [1.0, 2.0, 1.0, 2.0], []),                        # upper wall
        (9, [1], [1.0, 2.0, 1.0, 2.0], []),                        # lower wall
        (11, [0, 1, 2], [2.5, 1.0, 1.0, 1.0], [10]),               # OPES_EXPANDED
        (12, [0, 1], [], []),                                      # extended Lagrangian, D = 2
        (14, [2], [], []),                                         # EDS, D = 1
        (13, [0, 1, 2], [], []),                                   # MaxEnt, D = 3
        (15, [3], [2.5, 0.1, 0.2], [10]),                          # multithermal
        (101, [0, 1], [1.0, 0.1, 0.1, 10.0, 2.5, 0.0, 0.0], [1, 50]),  # hill list, D = 2
    ])


def _full_blob(rng, header=True):
    """Every section in the writer's order (CommonSerialization.cpp:20-178),
    with the values the test checks."""
    v = {"c": rng.normal(size=(3, 2)), "s": rng.uniform(0.05, 0.2, size=(3, 2)), "lw": rng.normal(size=3),
         "mean": rng.normal(size=2), "m2": rng.uniform(size=2), "rho": rng.normal(size=1),
         "g": rng.normal(size=16), "pb0": rng.normal(size=6), "pb1": rng.normal(size=6),
         "el_s": rng.normal(size=2), "el_p": rng.normal(size=2), "eds": rng.normal(size=4),
         "mx": rng.normal(size=3), "dF": rng.normal(size=2), "hc": rng.normal(size=(4, 2)),
         "hh": rng.uniform(size=4)}
    b = (b"GPUS" + i32(1)) if header else b""
    b += i32(1) + i32(3) + f64(0.75) + i32(12) + f64(*v["mean"]) + f64(*v["m2"])
    b += f64(*v["c"].reshape(-1)) + f64(*v["s"].reshape(-1)) + f64(*v["lw"])
    b += i32(1) + f64(*v["rho"])
    b += i32(1) + i32(9) + f64(*v["g"])
    b += i32(1) + i32(2) + i32(4) + f64(*v["pb0"]) + i32(5) + f64(*v["pb1"])
    b += i32(0)                      # external
    b += i32(1) + i32(2)             # linear, walls
    b += i32(1) + f64(-1.25) + i32(7)
    b += i32(1) + f64(*v["el_s"]) + f64(*v["el_p"])
    b += i32(1) + f64(v["eds"][0]) + f64(v["eds"][1]) + f64(v["eds"][2]) + f64(v["eds"][3]) + i32(17)
    b += i32(1) + f64(*v["mx"])
    b += i32(1) + i32(2) + f64(*v["dF"]) + f64(0.5) + f64(40.0)
    b += b"HILL" + i32(1) + i32(4) + i32(2) + f64(*v["hc"].reshape(-1)) + f64(*v["hh"])
    return b, v


def test_every_section_parsed_field_by_field():
    rng = np.random.default_rng(3)
    blob, v = _full_blob(rng)
    st = bs.parse(blob, _full_force())
    o = st.opes[0]
    assert (o.nk, o.sum_uprob, o.n_samples) == (3, 0.75, 12)
    for got, want in ((o.centers, v["c"]), (o.sigmas, v["s"]), (o.logw, v["lw"]), (o.mean, v["mean"]),
                      (o.m2, v["m2"]), (st.abmd[0], v["rho"]), (st.metad[0][1], v["g"]),
                      (st.pbmetad[0][0][1], v["pb0"]), (st.pbmetad[0][1][1], v["pb1"]),
                      (st.ext_lag[0][0], v["el_s"]), (st.ext_lag[0][1], v["el_p"]), (st.maxent[0], v["mx"]),
                      (st.multithermal[0][0], v["dF"]), (st.hills[0][0], v["hc"]), (st.hills[0][1], v["hh"])):
        assert np.array_equal(got, want)
    assert st.metad[0][0] == 9 and [s[0] for s in st.pbmetad[0]] == [4, 5]
    assert (st.n_external, st.n_linear, st.n_wall) == (0, 1, 2)
    assert st.opes_expanded == [(-1.25, 7)]
    assert [float(x[0]) for x in st.eds[0][:4]] == list(v["eds"]) and int(st.eds[0][4][0]) == 17
    assert st.multithermal[0][1:] == (0.5, 40.0)
    assert bs.pack(st) == blob


def test_repack_of_blobs_ending_early_and_legacy():
    """The reader stops where the bytes stop (CommonSerialization.cpp
    :180-410): a blob cut after any section, or without the header, packs
    back to the same bytes."""
    rng = np.random.default_rng(5)
    force = _full_force()
    full, _ = _full_blob(rng)
    st = bs.parse(full, force)
    cuts = []
    for k in range(len(bs.SECTIONS) + 1):
        st.n_sections = k
        cuts.append(bs.pack(st))
    assert len(set(len(c) for c in cuts)) == len(cuts)
    for k, c in enumerate(cuts):
        again = bs.parse(c, force)
        assert again.n_sections == k
        assert bs.pack(again) == c
    legacy, _ = _full_blob(np.random.default_rng(5), header=False)
    st = bs.parse(legacy, force)
    assert not st.header and bs.pack(st) == legacy == full[8:]


def test_metad_grid_size_follows_the_kernel_layout():
    """MetaD ints are [pace, bins..., periodic...] (CommonBiasKernels.cpp:206-233):
    bins (4, 3), periodic (1, 0) is 4 x 4 points, not the interleaved reading's 5 x 3."""
    f = Force(temperature=300.0)
    a = f.add_distance([0, 1])
    b = f.add_distance([0, 2])
    f.add_metad([a, b], 0.1, 1.0, 10, grid_min=[0.0, 0.0], grid_max=[1.0, 1.0], bins=[4, 3], periodic=[True, False])
    assert bs.ForceLayout(f).metad == [16]
    _, _, _, ip = f.getBiasInfo(0)
    assert ip == [10, 4, 3, 1, 0]


def test_mismatches_raise():
    blob = metad_blob([0.0] * 6)
    with pytest.raises(bs.BlobError):
        bs.parse(blob, _Layout([]))                                 # the force has no MetaD
    with pytest.raises(bs.BlobError):
        bs.parse(blob[:-6], META1)                                  # truncated inside a count
    with pytest.raises(bs.BlobError):
        bs.parse(blob + b"\0\0", META1)                             # trailing bytes
    with pytest.raises(bs.BlobError):
        bs.parse(b"GPUS" + i32(2) + blob[8:], META1)                # version


def test_incremental_merge_adds_only_each_groups_change():
    """new = shared + sum_g (G_g - base_g) (MultiGPUManager.py:208-215)."""
    base = [1.0] * 6
    first, baselines = bs.BiasStateMerger.merge_additive_incremental(
        [metad_blob(base, 2), metad_blob(base, 2)], META1, None)
    np.testing.assert_allclose(bs.parse(first, META1).metad[0][1], np.full(6, 2.0))
    # each group adds its own deposits to the shared grid it received
    g0, g1 = np.full(6, 2.0) + 0.25, np.full(6, 2.0) + np.arange(6.0)
    merged, _ = bs.BiasStateMerger.merge_additive_incremental(
        [metad_blob(list(g0), 5), metad_blob(list(g1), 7)], META1, baselines)
    st = bs.parse(merged, META1)
    np.testing.assert_allclose(st.metad[0][1], 2.0 + 0.25 + np.arange(6.0), rtol=1e-12)
    assert st.metad[0][0] == 4 + (5 - 4) + (7 - 4)


def _opes_blob(c, s, lw, su, ns, mean, m2):
    b = b"GPUS" + i32(1) + i32(1) + i32(len(lw)) + f64(su) + i32(ns) + f64(*mean) + f64(*m2)
    return b + f64(*c.reshape(-1)) + f64(*s.reshape(-1)) + f64(*lw)


def _welford(x):
    mean = x.mean(axis=0)
    return len(x), mean, ((x - mean) ** 2).sum(axis=0)


def test_opes_union_merge():
    """Kernel lists concatenated in blob order; Welford statistics equal to
    those of the pooled samples (Chan's combination is exact); the table's
    scalar field as the reference re-forms it, or the union's pairwise
    sum_uprob on request."""
    rng = np.random.default_rng(11)
    force = _Layout([(5, [0, 1], [2.5, 10.0, 0.1, 0.1, 0.01], [0, 1, 100])])
    xs = [rng.normal(size=(n, 2)) * [0.3, 0.1] + [1.0, 2.0] for n in (40, 25, 70)]
    blobs, tabs = [], []
    for x in xs:
        n = int(rng.integers(3, 9))
        c, s, lw = rng.normal(size=(n, 2)), rng.uniform(0.05, 0.2, size=(n, 2)), rng.normal(size=n)
        ns, mean, m2 = _welford(x)
        blobs.append(_opes_blob(c, s, lw, 1.5, ns, mean, m2))
        tabs.append((c, s, lw))
    st = bs.parse(bs.BiasStateMerger.merge_additive(blobs, force), force)
    o = st.opes[0]
    c = np.concatenate([t[0] for t in tabs])
    s = np.concatenate([t[1] for t in tabs])
    lw = np.concatenate([t[2] for t in tabs])
    assert np.array_equal(o.centers, c) and np.array_equal(o.sigmas, s) and np.array_equal(o.logw, lw)
    n, mean, m2 = _welford(np.concatenate(xs))
    assert o.n_samples == n
    np.testing.assert_allclose(o.mean, mean, rtol=1e-13)
    np.testing.assert_allclose(o.m2, m2, rtol=1e-12)
    assert o.sum_uprob == pytest.approx(math.log(np.sum(np.exp(lw))), rel=1e-14)
    st2 = bs.parse(bs.BiasStateMerger.merge_additive(blobs, force, opes_sum_uprob="recompute"), force)
    cut2 = 2.0 * 10.0 / ((10.0 - 1.0) / 10.0)
    brute = 0.0
    for j in range(len(lw)):
        for k in range(len(lw)):
            n2 = float(np.sum(((c[j] - c[k]) / s[k]) ** 2))
            if n2 < cut2:
                brute += math.exp(lw[k]) * (math.exp(-0.5 * n2) - math.exp(-0.5 * cut2))
    assert st2.opes[0].sum_uprob == pytest.approx(brute, rel=1e-12)


def test_opes_expanded_log_z_average():
    """log Z merged as log((n1 e^z1 + n2 e^z2) / (n1 + n2)); no updates: the plain mean."""
    force = _Layout([(11, [0, 1], [2.5, 1.0, 1.0], [10])])

    def blob(z, n):
        b = b"GPUS" + i32(1) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0)
        return b + i32(1) + f64(z) + i32(n)

    st = bs.parse(bs.BiasStateMerger.merge_additive([blob(-3.0, 4), blob(2.0, 12)], force), force)
    assert st.opes_expanded[0][1] == 16
    assert st.opes_expanded[0][0] == pytest.approx(math.log((4 * math.exp(-3.0) + 12 * math.exp(2.0)) / 16),
                                                   rel=1e-14)
    st = bs.parse(bs.BiasStateMerger.merge_additive([blob(-3.0, 0), blob(2.0, 0)], force), force)
    assert st.opes_expanded[0] == (-0.5, 0)


def test_hill_lists_concatenate_and_grow_incrementally():
    force = _Layout([(101, [0, 1], [1.0, 0.1, 0.1, 10.0, 2.5, 0.0, 0.0], [1, 50])])

    def blob(c, h):
        b = b"GPUS" + i32(1) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0) + i32(0)
        b += i32(0) + i32(0) + i32(0) + i32(0) + i32(0)
        return b + b"HILL" + i32(1) + i32(len(h)) + i32(2) + f64(*c.reshape(-1)) + f64(*h)

    rng = np.random.default_rng(2)
    c0, h0 = rng.normal(size=(3, 2)), rng.uniform(size=3)
    c1, h1 = rng.normal(size=(2, 2)), rng.uniform(size=2)
    merged, base = bs.BiasStateMerger.merge_additive_incremental([blob(c0, h0), blob(c1, h1)], force)
    st = bs.parse(merged, force)
    assert np.array_equal(st.hills[0][0], np.concatenate([c0, c1]))
    assert np.array_equal(st.hills[0][1], np.concatenate([h0, h1]))
    # both groups continue from the merged list and add one hill each
    sc, sh = st.hills[0]
    e0, e1 = rng.normal(size=(1, 2)), rng.normal(size=(1, 2))
    m2, _ = bs.BiasStateMerger.merge_additive_incremental(
        [blob(np.concatenate([sc, e0]), np.append(sh, 0.5)), blob(np.concatenate([sc, e1]), np.append(sh, 0.7))],
        force, base)
    st2 = bs.parse(m2, force)
    assert np.array_equal(st2.hills[0][0], np.concatenate([sc, e0, e1]))
    assert np.array_equal(st2.hills[0][1], np.append(sh, [0.5, 0.7]))
