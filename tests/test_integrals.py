"""Integral files: synthetic generator spec, FCIDUMP -> _IN converter against the
reference converter's own output (golden fixtures), and the loader."""
import filecmp
import os

import numpy as np
import pytest

import sq_bruteforce as B
from oracle import oracle as O
from sfci_amd import integrals as I

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("case", ["g3o2e", "g4o4e"])
def test_converter_matches_reference_bytes(tmp_path, case):
    out = tmp_path / (case + "_IN")
    I.convert_rhf(os.path.join(GOLDEN, case + "_fcidump"), str(out))
    assert filecmp.cmp(str(out), os.path.join(GOLDEN, case + "_IN"), shallow=False)


@pytest.mark.parametrize("case", ["u3o2e", "u4o4e"])
def test_uhf_converter_matches_reference_bytes(tmp_path, case):
    # the reference's conv_uhf (util/convert_fcidump.py:153-276, run with -u by
    # tests/golden/make_converter_golden.py) on synthetic spin-dependent integrals
    out = tmp_path / (case + "_IN")
    I.convert_uhf(os.path.join(GOLDEN, case + "_fcidump"), str(out))
    assert filecmp.cmp(str(out), os.path.join(GOLDEN, case + "_IN"), shallow=False)


def test_synthetic_generator_spec():
    h, canon, ec = I.synth_integrals(12, 0)
    n = 12
    assert ec == 0.0
    assert np.allclose(np.diag(h), np.linspace(-2.0, 1.5, n))
    assert np.array_equal(h, h.T)
    assert len(canon) == (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2
    for v, i, j, k, l in canon:
        assert i >= j and k >= l and i * (i + 1) // 2 + j >= k * (k + 1) // 2 + l
        if i == j and k == l:
            assert 0.3 <= v <= 0.8
    # deterministic
    h2, canon2, _ = I.synth_integrals(12, 0)
    assert np.array_equal(h, h2) and canon == canon2
    h3, _, _ = I.synth_integrals(12, 1)
    assert not np.array_equal(h, h3)
    h4, _, _ = I.synth_integrals(50, 3, eps_lo=-1.0, eps_hi=0.5, sigma_g=0.04)
    assert np.isclose(h4[0, 0], -1.0) and np.isclose(h4[-1, -1], 0.5)


def test_fcidump_roundtrip(tmp_path):
    h, canon, ec = I.synth_integrals(5, 3)
    p = str(tmp_path / "x_fcidump")
    I.write_fcidump(p, h, canon, 4, ec)
    h2, eri, ec2, n = I.read_fcidump(p)
    assert n == 5 and ec2 == ec and np.array_equal(h2, h)
    for v, i, j, k, l in canon:
        for a, b, c, d in ((i, j, k, l), (j, i, l, k), (k, l, i, j), (l, k, j, i)):
            assert eri[a, b, c, d] == v


def test_in_file_entry_counts(make_system):
    # every spin-allowed (mn|pq) with m != p and n != q (same spin), all permutations once
    s = make_system(3, 2, 0)
    lines = open(s["in_path"]).read().split("\n")
    i, blocks = 0, []
    while i < len(lines) and lines[i].strip():
        c = int(lines[i])
        blocks.append(lines[i + 1:i + 1 + c])
        i += 1 + c
    assert len(blocks) == 7
    n = 3
    # one-body: n*(n+1)/2 lines, off-diagonal doubled, per spin
    assert len(blocks[4]) == len(blocks[5]) == n * n
    assert len(blocks[6]) == 1
    entries = set()
    for b in blocks[:4]:
        for line in b:
            t = line.split()
            entries.add(tuple(int(x) for x in t[1:]))
    nso = 2 * n
    expect = 0
    for m in range(nso):
        for nn in range(nso):
            for p in range(nso):
                for q in range(nso):
                    if (m & 1) != (nn & 1) or (p & 1) != (q & 1):
                        continue
                    if (m & 1) == (p & 1) and (m == p or nn == q):
                        continue
                    expect += 1
    assert len(entries) == expect


def test_real_molecule_c2_hf_energy(tmp_path):
    """The reference's C2/cc-pVDZ example (symmetry-sparse FCIDUMP, 26 orbitals, 8
    correlated electrons; example/C2, run by example/sfci.input with
    init_hf=202125315): the converted _IN gives E_HF equal to the closed form from the
    FCIDUMP's spatial integrals."""
    from oracle import oracle as O
    fc = os.path.join(GOLDEN, "c2-cc-pvdz-hf_fcidump")
    inp = I.convert_rhf(fc, str(tmp_path / "c2_IN"))
    o = O.Oracle(inp, 52, 8)
    hf = 202125315
    occ = [i // 2 for i in range(52) if (hf >> i) & 1 and i % 2 == 0]
    assert occ == [0, 6, 9, 13]
    h, eri, ec, n = I.read_fcidump(fc)
    e_hf = ec + 2 * sum(h[i, i] for i in occ) + sum(2 * eri[i, i, j, j] - eri[i, j, j, i] for i in occ for j in occ)
    assert abs(o.diag(hf) - e_hf) < 1e-10
    # symmetry-sparse: many connections have an exactly zero element
    k, v = o.connections(hf)
    assert (v == 0.0).sum() > 0 and len(k) == 1 + 2 * 4 * 22 + 2 * 6 * 231 + (4 * 22) ** 2


def _uhf_system(tmp_path, n_orb, n_elec, seed, same_spin=False):
    h_a, h_b, g_aa, g_bb, g_ab, ec = I.synth_uhf_integrals(n_orb, seed)
    if same_spin:
        h_b, g_aa, g_bb, g_ab = h_a, g_ab, g_ab, g_ab
        g_ab = 0.5 * (g_ab + g_ab.transpose(2, 3, 0, 1))
        g_aa = g_bb = g_ab
    fc = str(tmp_path / ("uhf%d_fcidump" % n_orb))
    I.write_fcidump_uhf(fc, h_a, h_b, g_aa, g_bb, g_ab, n_elec, ec)
    inp = I.convert_uhf(fc, str(tmp_path / ("uhf%d_IN" % n_orb)))
    return inp, (h_a, h_b, g_aa, g_bb, g_ab, ec)


def test_convert_uhf_block_layout(tmp_path):
    inp, _ = _uhf_system(tmp_path, 4, 4, 2)
    lines = open(inp).read().split("\n")
    counts, t = [], 0
    while t < len(lines) and lines[t].strip():
        c = int(lines[t])
        counts.append(c)
        blk = lines[t + 1:t + 1 + c]
        for ln in blk:
            v, m, n, p, q = ln.split()
            m, n, p, q = int(m), int(n), int(p), int(q)
            if len(counts) <= 2:
                sp = {1: 1, 2: 0}[len(counts)]
                assert (m & 1, n & 1, p & 1, q & 1) == (sp,) * 4
            elif len(counts) <= 4:
                # opposite-spin blocks hold both (ab|..) and their (..|ab) images
                assert (m & 1, n & 1, p & 1, q & 1) in ((1, 1, 0, 0), (0, 0, 1, 1))
        t += 1 + c
    assert len(counts) == 7 and counts[6] == 1
    # every aa / bb / ab spin-orbital key appears once; the du block adds nothing that the
    # ud block's (p, q, m, n) images did not already hold
    assert counts[3] == 0
    n = 4
    assert counts[4] == counts[5] == n * n


def test_convert_uhf_equals_rhf_for_spin_free_integrals(tmp_path, make_system):
    # the same spatial integrals written as an RHF and as a UHF FCIDUMP give the same H
    h, canon, _ = I.synth_integrals(6, 4)
    n = 6
    g = np.zeros((n,) * 4)
    for v, i, j, k, l in canon:
        for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                           (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
            g[a, b, c, d] = v
    fr = str(tmp_path / "r_fcidump")
    I.write_fcidump(fr, h, canon, 6, 1.25)
    ir = I.convert_rhf(fr, str(tmp_path / "r_IN"))
    fu = str(tmp_path / "u_fcidump")
    I.write_fcidump_uhf(fu, h, h, g, g, g, 6, 1.25)
    iu = I.convert_uhf(fu, str(tmp_path / "u_IN"))
    orr, ou = O.Oracle(ir, 12, 6), O.Oracle(iu, 12, 6)
    dets = np.array(B.sz0_dets(6, 3), dtype=np.uint64)
    rng = np.random.default_rng(0)
    k = np.sort(rng.choice(dets, size=150, replace=False))
    v = rng.normal(size=len(k))
    a, b = orr.hv(k, v, cut=0.0), ou.hv(k, v, cut=0.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_uhf_slater_condon_vs_second_quantized(tmp_path):
    inp, (h_a, h_b, g_aa, g_bb, g_ab, ec) = _uhf_system(tmp_path, 5, 4, 3)
    o = O.Oracle(inp, 10, 4)
    hd = {1: h_a, 0: h_b}
    ed = {(1, 1): g_aa, (0, 0): g_bb, (1, 0): g_ab, (0, 1): g_ab.transpose(2, 3, 0, 1)}
    dets = B.sz0_dets(5, 2)
    for d in dets[::7]:
        ref = B.apply_h(hd, ed, ec, d, 5)
        k, v = o.connections(d)
        got = dict(zip(k.tolist(), v.tolist()))
        for t, x in ref.items():
            assert abs(got.get(t, 0.0) - x) < 1e-12, (d, t)
        for t, x in got.items():
            assert abs(ref.get(t, 0.0) - x) < 1e-12, (d, t)


def test_workload_table():
    """The benchmark workloads: orbital space, electrons, seed, Nd, Nl, restarts and the
    generator settings of each synthetic system."""
    W = I.WORKLOADS
    assert W["c1"] == dict(n_orb=12, n_elec=8, seed=0, nd=20000, nl=20, restarts=5, gen={})
    assert W["c2"] == dict(n_orb=26, n_elec=14, seed=1, nd=1_000_000, nl=30, restarts=8, gen={})
    assert W["c3"] == dict(n_orb=32, n_elec=16, seed=2, nd=4_000_000, nl=30, restarts=8, gen={})
    assert W["c4"] == dict(n_orb=50, n_elec=20, seed=3, nd=10_000_000, nl=40, restarts=10,
                           gen=dict(eps_lo=-1.0, eps_hi=0.5, sigma_g=0.04))
    assert W["c5"] == dict(n_orb=32, n_elec=16, seed=4, nd=4_000_000, nl=30, restarts=8, gen={})
    # the synthetic generator's defaults: orbital energies from -2.0 to +1.5 Ha, h off-diagonal
    # N(0, 0.05), (pq|rs) off-diagonal N(0, 0.02)
    import inspect
    sig = inspect.signature(I.synth_integrals).parameters
    assert (sig["eps_lo"].default, sig["eps_hi"].default, sig["sigma_h"].default, sig["sigma_g"].default) == \
        (-2.0, 1.5, 0.05, 0.02)
