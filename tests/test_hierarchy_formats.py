"""Hierarchy oracle pinned (reference test, closed-form eigenvectors, an
independent recursion), hierarchy shape and regions (host code of the
library), BLIF reader, DEF writer and host HPWL. No GPU."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import fluxion as fxn
from fluxion import _lib
from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def chain(n):
    return fxn.Circuit.from_nets(n, [[i, i + 1] for i in range(n - 1)])


# ---- oracle: build_hierarchy ----

@pytest.mark.parametrize("mode", [0, 2])
def test_oracle_partition_small_chain(mode):
    # partitioner.rs:232-245 (test_partition_small_chain): chain of 20 gates,
    # max_leaf_size 8: depth >= 1, >= 2 leaves, every gate in a leaf.
    c = chain(20)
    aptr, adj = c.build_adjacency()
    order, nodes = orc.build_hierarchy(20, aptr, adj, max_leaf=8, mode=mode)
    leaves = np.flatnonzero(nodes["left"] < 0)
    assert nodes["depth"].max() >= 1 and len(leaves) >= 2
    assert sorted(order.tolist()) == list(range(20))
    assert nodes["count"][leaves].max() <= 8


def laplacian(n, aptr, adj):
    # L = D - A over adjacency entries, as partitioner.rs:128-137 counts them.
    L = np.zeros((n, n))
    for i in range(n):
        for k in range(aptr[i], aptr[i + 1]):
            L[i, i] += 1.0
            L[i, adj[k]] -= 1.0
    return L


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("graph", ["chain12", "random40"])
def test_oracle_power_iteration_converges_to_top_eigenvector(mode, graph):
    # compute_fiedler_vector is plain power iteration on L with the constant
    # vector deflated: with enough iterations it reaches L's eigenvector of
    # largest eigenvalue. numpy's symmetric eigensolver gives that vector.
    if graph == "chain12":
        c = chain(12)
        # closed form for a path: lambda_k = 2 - 2 cos(pi k / n)
        lam = 2 - 2 * np.cos(np.pi * np.arange(12) / 12)
    else:
        rng = np.random.default_rng(11)
        nets = [sorted(set(rng.integers(0, 40, rng.integers(2, 5)).tolist())) for _ in range(60)]
        c = fxn.Circuit.from_nets(40, [n for n in nets if len(n) >= 2])
        lam = None
    aptr, adj = c.build_adjacency()
    n = c.num_gates
    w, V = np.linalg.eigh(laplacian(n, aptr, adj))
    if lam is not None:
        assert np.allclose(np.sort(lam), w, atol=1e-9)
    top = V[:, -1]
    assert w[-1] - w[-2] > 1e-3
    v = orc.fiedler_vector(np.arange(n), n, aptr, adj, iters=4000, mode=mode).astype(np.float64)
    s = np.sign(v @ top)
    assert np.abs(v - s * top).max() < 2e-5
    # and a dropped degree term or a flipped sign would not converge there
    assert abs(np.linalg.norm(v) - 1.0) < 1e-5


def test_oracle_fiedler_matches_power_iter_restatement():
    c, _, _ = fxn.netgen.flat(1500, 1650, 7)
    aptr, adj = c.build_adjacency()
    gates = np.random.default_rng(5).permutation(1500)[:900].astype(np.int32)
    a = orc.fiedler_vector(gates, 1500, aptr, adj, iters=30, mode=0)
    b = orc.power_iter(gates, 1500, aptr, adj, fxn.fiedler_init(900), 30, mode=0)
    assert np.array_equal(a, b)
    d = orc.fiedler_vector(gates, 1500, aptr, adj, iters=30, mode=2)
    assert np.abs(d - a).max() < 1e-5


def _py_tree(n_total, aptr, adj, gates, cfg, iters, mode, nodes, order, start, depth):
    # recursive_bisect (partitioner.rs:64-100) in Python on top of the
    # oracle's Fiedler vector, numpy's stable argsort and the split rule.
    mx, mn, bf = cfg
    idx = len(nodes)
    nodes.append([start, len(gates), depth, -1, -1])
    k = len(gates)
    if k <= mx or k <= mn:
        order[start:start + k] = gates
        return idx
    f = orc.fiedler_vector(gates, n_total, aptr, adj, iters, mode)
    f = np.where(f == 0, np.float32(0), f)
    o = np.argsort(f, kind="stable")
    ms = int(math.ceil(np.float32(k) * np.float32(bf)))
    split = min(max(k // 2, ms), k - ms)
    if split <= 0 or split >= k:
        order[start:start + k] = gates
        return idx
    g = np.asarray(gates)[o]
    nodes[idx][3] = _py_tree(n_total, aptr, adj, g[:split], cfg, iters, mode, nodes, order, start,
                             depth + 1)
    nodes[idx][4] = _py_tree(n_total, aptr, adj, g[split:], cfg, iters, mode, nodes, order,
                             start + split, depth + 1)
    return idx


@pytest.mark.parametrize("mode", [0, 2])
def test_oracle_hierarchy_matches_independent_recursion(mode):
    c, _, _ = fxn.netgen.flat(3000, 3300, 5)
    aptr, adj = c.build_adjacency()
    order, nodes = orc.build_hierarchy(3000, aptr, adj, max_leaf=40, mode=mode)
    pn, po = [], np.zeros(3000, np.int32)
    _py_tree(3000, aptr, adj, np.arange(3000, dtype=np.int32), (40, 8, 0.4), 30, mode, pn, po, 0, 0)
    pn = np.array(pn)
    assert np.array_equal(po, order)
    for i, k in enumerate(("start", "count", "depth", "left", "right")):
        assert np.array_equal(pn[:, i], nodes[k])


# ---- library host code: hierarchy shape and regions ----

@pytest.mark.parametrize("n,cfg,want", [
    (20, (8, 8, 0.4), (7, 2)),         # 20 -> 10+10 -> 5+5 each
    (5, (1, 0, 0.5), (9, 3)),          # 5 -> 2+3, 2 -> 1+1, 3 -> 1+2, 2 -> 1+1
    (1000, (64, 8, 1.0), (1, 0)),      # balance 1.0 empties a side: the root is a leaf
    (0, (64, 8, 0.4), (1, 0)),
    (64, (64, 8, 0.4), (1, 0)),
    (65, (64, 8, 0.4), (3, 1)),
])
def test_hierarchy_size_kats(n, cfg, want):
    pc = fxn.PartitionConfig(*cfg)
    assert fxn.hierarchy_size(n, pc) == want
    if n:
        _, nodes = orc.build_hierarchy(n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32),
                                       max_leaf=cfg[0], min_cluster=cfg[1], balance=cfg[2], iters=2)
        assert len(nodes["start"]) == want[0] and nodes["depth"].max() == want[1]


def test_hierarchy_config_rejected():
    for bad in (fxn.PartitionConfig(balance_factor=1.5), fxn.PartitionConfig(balance_factor=-0.1),
                fxn.PartitionConfig(max_leaf_size=-1), fxn.PartitionConfig(power_iterations=-2)):
        with pytest.raises(fxn.FluxionError):
            fxn.hierarchy_size(100, bad)


def lib_regions(nodes, w, h):
    nn = len(nodes["left"])
    out = np.zeros((nn, 4), np.float32)
    left = np.ascontiguousarray(nodes["left"], np.int32)
    right = np.ascontiguousarray(nodes["right"], np.int32)
    _lib.check(_lib.lib.fx_hierarchy_regions(nn, _lib.ptr(left), _lib.ptr(right), w, h,
                                             _lib.ptr(out)))
    return out


def test_regions_kat_chain20():
    c = chain(20)
    aptr, adj = c.build_adjacency()
    _, nodes = orc.build_hierarchy(20, aptr, adj, max_leaf=8)
    # ids: 0 root, 1 L, 2 LL, 3 LR, 4 R, 5 RL, 6 RR; every split halves the
    # width (two children: cols = ceil(sqrt(2)) = 2, rows = 1).
    want = np.array([[500, 500, 1000, 1000], [250, 500, 500, 1000], [125, 500, 250, 1000],
                     [375, 500, 250, 1000], [750, 500, 500, 1000], [625, 500, 250, 1000],
                     [875, 500, 250, 1000]], np.float32)
    assert np.array_equal(lib_regions(nodes, 1000.0, 1000.0), want)
    assert np.array_equal(orc.hapr_regions(nodes, 1000.0, 1000.0), want)


def test_regions_lib_matches_oracle():
    c, _, _ = fxn.netgen.flat(5000, 5500, 9)
    aptr, adj = c.build_adjacency()
    _, nodes = orc.build_hierarchy(5000, aptr, adj, max_leaf=30, mode=2, iters=5)
    a = lib_regions(nodes, 777.7, 333.3)
    b = orc.hapr_regions(nodes, 777.7, 333.3)
    assert np.array_equal(a, b)


# ---- BLIF reader ----

def test_blif_reference_fixture():
    # HAPR_fluxion_v4/tests/fixtures/small.blif: 4 gates, 2 multi-gate nets.
    c = fxn.parse_blif(os.path.join(GOLDEN, "small.blif"))
    assert c.name == "small" and c.num_gates == 4 and c.num_nets == 2
    nets = {c.net_names[e]: c.net_pins[c.net_ptr[e]:c.net_ptr[e + 1]].tolist()
            for e in range(c.num_nets)}
    # n1: driven by g0 (AND2), read by g2 (XOR2); n2: g1 (OR2) -> g2, g3 (INV)
    assert nets == {"n1": [0, 2], "n2": [1, 2, 3]}
    assert c.chip_width == 1000.0 and c.chip_height == 1000.0
    aptr, adj = c.build_adjacency()
    assert [adj[aptr[g]:aptr[g + 1]].tolist() for g in range(4)] == [[2], [2, 3], [0, 1, 3], [1, 2]]


def test_blif_semantics(tmp_path):
    text = """# comment
.model t
.inputs a b c
.names a b x
11 1
.names x c y

.latch y q re clk 0
.gate NAND2 A=q B=a Y=z
.gate BUF A=z O=x2 O=w
.names a z x2
.names q
.names
.end
"""
    f = tmp_path / "t.v1.blif"
    f.write_text(text)
    c = fxn.parse_blif(f)
    assert c.name == "t.v1"
    # g0 .names a b x | g1 .names x c y | g2 .latch y q | g3 .gate NAND2 |
    # g4 .gate BUF | g5 .names a z x2 | g6 .names q (drives q, no inputs) | g7 .names
    assert c.num_gates == 8
    assert c.gate_types.tolist() == [0, 0, 1, 0, 0, 0, 0, 0]
    nets = {c.net_names[e]: c.net_pins[c.net_ptr[e]:c.net_ptr[e + 1]].tolist()
            for e in range(c.num_nets)}
    # x: g0 -> g1; y: g1 -> g2; q: last driver g6 (replaces g2), read by g3;
    # z: g3 -> g4, g5; x2: last driver g5 (replaces g4), no readers -> dropped;
    # w: g4, no readers -> dropped; a: read by g0, g3, g5 (no driver);
    # b, c: one reader each -> dropped.
    assert nets == {"x": [0, 1], "y": [1, 2], "q": [6, 3], "z": [3, 4, 5], "a": [0, 3, 5]}
    # driven nets first (first-drive order), then readers-only nets
    assert c.net_names == ["x", "y", "q", "z", "a"]


def test_circuit_net_names_buffer():
    # fx_circuit_net_names: size query, exact-size copy, and a short buffer.
    import ctypes as C
    from fluxion._lib import check, lib
    h = C.c_void_p()
    check(lib.fx_blif_read(os.path.join(GOLDEN, "small.blif").encode(), C.byref(h)))
    try:
        need = C.c_int64()
        assert lib.fx_circuit_net_names(h, None, 0, C.byref(need)) == 0
        assert need.value == len(b"n1\0n2\0")
        buf = C.create_string_buffer(need.value)
        assert lib.fx_circuit_net_names(h, buf, need.value, C.byref(need)) == 0
        assert buf.raw == b"n1\0n2\0"
        assert lib.fx_circuit_net_names(h, buf, need.value - 1, C.byref(need)) != 0
    finally:
        lib.fx_circuit_free(h)


def test_blif_errors(tmp_path):
    with pytest.raises(fxn.FluxionError, match="Failed to open BLIF"):
        fxn.parse_blif(tmp_path / "missing.blif")
    f = tmp_path / "bad.blif"
    f.write_text(".gate\n")
    with pytest.raises(fxn.FluxionError, match=".gate without a gate type"):
        fxn.parse_blif(f)
    e = tmp_path / "empty.blif"
    e.write_text(".model e\n.end\n")
    c = fxn.parse_blif(e)
    assert c.num_gates == 0 and c.num_nets == 0


# ---- DEF writer ----

def test_def_writer_kat(tmp_path):
    c = fxn.Circuit.from_nets(4, [[0, 1]], 123.4567, 50.0)
    c.name = "kat"
    x = np.array([1.2345, -0.0015, np.nan, 1e30], np.float32)
    y = np.array([0.0, 999.9999, -1e30, 2.5], np.float32)
    pl = fxn.Placement(x, y)
    out = tmp_path / "o.def"
    fxn.write_def(out, c, pl)
    # f32 products truncated toward zero (Rust `as i64`): 1.2345f*1000 =
    # 1234.5 -> 1234; -1.5 -> -1; NaN -> 0; 1e33 and -1e33 saturate;
    # 999.9999f*1000 = 999999.9 -> 999999; 123.4567f*1000 = 123456.7 -> 123456.
    want = ("VERSION 5.8 ;\nDIVIDERCHAR \"/\" ;\nBUSBITCHARS \"[]\" ;\nDESIGN kat ;\n"
            "UNITS DISTANCE MICRONS 1000 ;\n\nDIEAREA ( 0 0 ) ( 123456 50000 ) ;\n\n"
            "COMPONENTS 4 ;\n"
            "    - g0 GATE_TYPE + PLACED ( 1234 0 ) N ;\n"
            "    - g1 GATE_TYPE + PLACED ( -1 999999 ) N ;\n"
            "    - g2 GATE_TYPE + PLACED ( 0 -9223372036854775808 ) N ;\n"
            "    - g3 GATE_TYPE + PLACED ( 9223372036854775807 2500 ) N ;\n"
            "END COMPONENTS ;\n\nEND DESIGN\n")
    assert out.read_text() == want
    c.gate_names = ["AND_1", "b", "c", "d"]
    fxn.write_def(out, c, pl)
    assert "    - AND_1 GATE_TYPE + PLACED ( 1234 0 ) N ;\n" in out.read_text()


def test_def_writer_integration_checks(tmp_path):
    # HAPR_fluxion_v4/tests/test_integration.py:30-38 on the BLIF fixture.
    c = fxn.parse_blif(os.path.join(GOLDEN, "small.blif"))
    pl = fxn.Placement(np.linspace(0, 900, 4).astype(np.float32), np.full(4, 10.0, np.float32))
    out = tmp_path / "small.def"
    fxn.write_def(out, c, pl)
    t = out.read_text()
    assert "COMPONENTS" in t and "END COMPONENTS" in t and "+ PLACED" in t
    assert "DESIGN small ;" in t and t.count("GATE_TYPE") == 4


# ---- host HPWL (placement.rs:48-80) ----

def test_host_hpwl():
    # placement.rs:86-97 (test_hpwl): one 4-pin net on a 10x10 square -> 20.
    pl = fxn.Placement(np.array([0, 10, 10, 0], np.float32), np.array([0, 0, 10, 10], np.float32))
    assert pl.compute_hpwl(fxn.Circuit.from_nets(4, [[0, 1, 2, 3]])) == 20.0
    # single-pin and empty nets are skipped; f32 accumulation in net order
    circ, x, y = fxn.netgen.flat(5000, 5500, 3)
    nets = [circ.net_pins[circ.net_ptr[e]:circ.net_ptr[e + 1]].tolist() for e in range(200)]
    nets = [[]] + nets[:50] + [[7]] + nets[50:] + [[]]
    c2 = fxn.Circuit.from_nets(5000, nets)
    pl = fxn.Placement(x, y)
    assert pl.compute_hpwl(c2) == orc.hpwl(c2.net_ptr, c2.net_pins, x, y, dtype=np.float32)


def test_engine_config(tmp_path):
    eng = fxn.FluxionEngine(os.path.join(GOLDEN, "default.yaml"))
    assert eng.partition_config() == fxn.PartitionConfig(64, 8, 0.4, 30)
    pc = eng.placer_config()
    assert (pc.force_iterations, pc.attract_weight, pc.repel_weight, pc.damping, pc.dt) == \
        (50, 1.0, 0.5, 0.9, 0.1)
    assert eng.refine_config() == fxn.RefineConfig(100, 5.0, 1.0, 256, 256)
    with pytest.raises(FileNotFoundError):
        fxn.FluxionEngine(tmp_path / "nope.yaml")
    with pytest.raises(FileNotFoundError):
        eng.place(tmp_path / "nope.blif", tmp_path / "o.def")


# ---- oracle: attention (attention.rs:50-88) ----

def test_oracle_attention_reference_case():
    # attention.rs:122-143 (test_attention_basic): 6 gates, nets n0 = [0, 1],
    # n1 = [0, 3], n2 = [3, 4]; clusters {0, 1, 2} and {3, 4, 5}. Each cluster
    # touches two nets and shares n1: scores 1/2 both ways.
    c = fxn.Circuit.from_nets(6, [[0, 1], [0, 3], [3, 4]], 100.0, 100.0)
    scores, top = orc.attention_level(6, c.net_ptr, c.net_pins, [0, 3], [3, 3], np.arange(6), 5)
    assert np.array_equal(scores, np.array([[0, 0.5], [0.5, 0]], np.float32))
    assert top.tolist() == [[1, -1, -1, -1, -1], [0, -1, -1, -1, -1]]


def test_oracle_attention_kat():
    # A = {0,1,2}, B = {3,4,5}, C = {6,7,8}, D = {9}; nets e0 = [0,3] (A,B),
    # e1 = [1,6] (A,C), e2 = [2] (single pin, A), e3 = [4,5] (B), e4 = [7,8] (C).
    # N_A = {e0,e1,e2}, N_B = {e0,e3}, N_C = {e1,e4}, N_D = {}.
    nets = [[0, 3], [1, 6], [2], [4, 5], [7, 8]]
    c = fxn.Circuit.from_nets(10, nets)
    order = np.arange(10)
    scores, top = orc.attention_level(10, c.net_ptr, c.net_pins, [0, 3, 6, 9], [3, 3, 3, 1], order, 3)
    third = np.float32(1) / np.float32(3)
    want = np.array([[0, third, third, 0], [0.5, 0, 0, 0], [0.5, 0, 0, 0], [0, 0, 0, 0]], np.float32)
    assert np.array_equal(scores, want)
    # ties keep ascending j; zero scores dropped; D has no nets: empty row
    assert top.tolist() == [[1, 2, -1], [0, -1, -1], [0, -1, -1], [-1, -1, -1]]
    _, top1 = orc.attention_level(10, c.net_ptr, c.net_pins, [0, 3, 6, 9], [3, 3, 3, 1], order, 1)
    assert top1.tolist() == [[1], [0], [0], [-1]]


def test_oracle_attention_symmetry():
    # shared counts are symmetric: score_ij * total_i == score_ji * total_j
    circ, _, _ = fxn.netgen.flat(600, 660, 4)
    starts = np.arange(0, 600, 50)
    counts = np.full(12, 50)
    scores, _ = orc.attention_level(600, circ.net_ptr, circ.net_pins, starts, counts,
                                    np.arange(600), 10)
    totals = []
    for s in starts:
        g = set(range(s, s + 50))
        totals.append(sum(1 for e in range(circ.num_nets)
                          if g & set(circ.net_pins[circ.net_ptr[e]:circ.net_ptr[e + 1]].tolist())))
    shared = scores.astype(np.float64) * np.array(totals)[:, None]
    assert np.allclose(shared, np.rint(shared), atol=1e-3)
    assert np.array_equal(np.rint(shared), np.rint(shared).T)
