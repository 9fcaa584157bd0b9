"""The C ABI of libparadmm without a GPU: the library loads, exports every
symbol the headers declare, builds graphs exactly as the reference does, and
its error paths return codes before any device work."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    names = set()
    for h in ("paradmm.h", "paradmm_gpu.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            names.add(m.group(1))
    drop = {"if", "defined", "sizeof", "op", "Operator", "paradmm_prox_fn", "void"}
    return sorted(n for n in names if n not in drop and not n.startswith("__"))


def test_library_exports_every_declared_symbol(pdm):
    out = subprocess.run(["nm", "-D", "--defined-only", pdm.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = header_symbols()
    assert len(declared) > 40
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    for n in pdm.exported_symbols():
        assert n in exported


def test_builder_matches_reference_structure(pdm, oracle):
    """Edge numbering, edge->var and per-variable lists equal the reference builder's."""
    g = pdm.packing_square(9, 1)
    o = oracle.Graph.square(9, 1)
    assert (g.nedge, g.nvar, g.nfac) == (o.nedge, o.nvar, o.nfac) == (90, 9, 9 + 9 + 36)
    assert np.array_equal(g.edge_var(), o.edge_var())
    lib = o.lib
    for v in range(g.nvar):
        deg = g.g.varix_to_numedges[v]
        assert deg == lib.or_var_deg(o.h, v) == 10
        mine = [g.g.varix_to_edgelistix[v][i] for i in range(deg)]
        theirs = [lib.or_var_edges(o.h, v)[i] for i in range(deg)]
        assert mine == theirs
    first = np.ctypeslib.as_array(lib.or_fac_first(o.h), shape=(o.nfac,))
    for f in range(g.nfac):
        assert g.g.operatorix_to_edgelistix[f][0] == first[f]
    for name, oname in [("X", "X"), ("N", "N"), ("M", "M"), ("U", "U"), ("Z", "Z"), ("RHOS", "RHO")]:
        assert np.array_equal(g.array(name), o.array(oname)), name


def test_addnode_sorts_a_copy_and_handles_growth(pdm):
    g = pdm.Graph(2)
    addr = pdm.prox_address("paradmm_prox_disk_collision_2d")
    rng = np.random.default_rng(0)
    lists = []
    for _ in range(3000):  # crosses several capacity doublings
        v = rng.choice(200, size=4, replace=False)
        arr = (C.c_int * 4)(*v.tolist())
        g.lib.addNode(g.ptr, C.c_void_p(addr), None, 4, arr)
        assert list(arr) == v.tolist()  # caller's list untouched
        lists.append(sorted(v.tolist()))
    assert g.nfac == 3000 and g.nedge == 12000
    ev = g.edge_var()
    for f in (0, 1, 777, 2999):
        s = g.g.operatorix_to_edgelistix[f][0]
        assert s == 4 * f
        assert [g.g.operatorix_to_edgelistix[f][i] for i in range(4)] == [s, s + 1, s + 2, s + 3]
        assert ev[s:s + 4].tolist() == lists[f]
    total = sum(g.g.varix_to_numedges[v] for v in range(g.nvar))
    assert total == 12000


def test_unknown_callback_is_reported_when_host_callbacks_are_off(pdm, monkeypatch):
    """With PARADMM_HOST_PROX=0 a callback the engine cannot run on the GPU is
    an error code before any device work; by default such a callback runs on
    the host (tested on the GPU)."""
    monkeypatch.setenv("PARADMM_HOST_PROX", "0")
    cb = pdm.PROX_FN(lambda *a: None)
    g = pdm.Graph(3)
    vv = (C.c_int * 1)(0)
    g.lib.addNode(g.ptr, C.cast(cb, C.c_void_p), None, 1, vv)
    g.init_rho_alpha(1.0, 1.0)
    g.init_zeros()
    rc = g.lib.paradmm_gpu_prepare(g.ptr)
    assert rc == 2
    assert "not a device prox kind" in g.lib.paradmm_last_error().decode()
    assert g.lib.paradmm_prox_kind_of(C.cast(cb, C.c_void_p)) == 0
    # registering it as a kind makes it acceptable (checked again on the GPU)
    assert g.lib.paradmm_register_prox_kind(C.cast(cb, C.c_void_p), 5) == 0
    assert g.lib.paradmm_prox_kind_of(C.cast(cb, C.c_void_p)) == 5
    g.close()


def test_shape_mismatch_is_reported(pdm):
    g = pdm.Graph(2)  # dim-3 operator on a dim-2 graph
    vv = (C.c_int * 2)(0, 1)
    g.lib.addNode(g.ptr, C.c_void_p(pdm.prox_address("paradmm_prox_disk_collision_3d")), None, 2, vv)
    g.init_rho_alpha(1.0, 1.0)
    g.init_zeros()
    assert g.lib.paradmm_gpu_prepare(g.ptr) == 3
    assert "do not fit" in g.lib.paradmm_last_error().decode()


def test_builtin_kinds_registered(pdm):
    lib = pdm.load()
    for name, k in [("paradmm_prox_disk_collision_2d", 1), ("paradmm_prox_disk_plane_2d", 2),
                    ("paradmm_prox_max_radius_2d", 3), ("paradmm_prox_disk_collision_3d", 4),
                    ("paradmm_prox_box_3d", 5), ("paradmm_prox_max_radius_3d", 6)]:
        assert lib.paradmm_prox_kind_of(C.c_void_p(pdm.prox_address(name))) == k


def test_utilities(pdm):
    lib = pdm.load()
    assert lib.abs_d(-2.5) == 2.5 and lib.abs_d(3.0) == 3.0
    a = np.random.default_rng(2).integers(-50, 50, size=300).astype(np.int32)
    arr = (C.c_int * len(a))(*a.tolist())
    lib.quicksort(arr, len(a))
    assert list(arr) == sorted(a.tolist())
    r = [lib.get_random(-1.0, 2.0) for _ in range(100)]
    assert min(r) >= -1.0 and max(r) <= 2.0


def test_seeded_init_equals_oracle(pdm, oracle):
    g = pdm.packing_square(30, 77, 2.0, 0.5)
    o = oracle.Graph.square(30, 77, 2.0, 0.5)
    assert np.array_equal(g.array("Z"), o.array("Z"))
    assert np.array_equal(g.array("X"), o.array("X"))
    assert np.array_equal(g.array("ALPHAS"), o.array("ALPHA"))


def test_add_nodes_matches_addnode_loop(pdm):
    """The bulk call numbers edges and builds adjacency exactly as the same
    addNode calls do."""
    rng = np.random.default_rng(4)
    pairs = np.array([(i, j) for i in range(12) for j in range(i + 1, 12)], dtype=np.int32)
    rng.shuffle(pairs)
    pairs[::3] = pairs[::3, ::-1]  # unsorted variable lists too
    fn = pdm.prox_address("paradmm_prox_disk_collision_3d")
    a, b = pdm.Graph(3), pdm.Graph(3)
    for i, j in pairs:
        a.add_factor(fn, [int(i), int(j)])
    b.add_factors(fn, pairs)
    assert a.g.num_opts == b.g.num_opts and a.g.num_edges == b.g.num_edges and a.g.num_vars == b.g.num_vars
    for e in range(a.g.num_edges):
        assert a.g.edgeix_to_varlistix[e] == b.g.edgeix_to_varlistix[e]
    for v in range(a.g.num_vars):
        n = a.g.varix_to_numedges[v]
        assert n == b.g.varix_to_numedges[v]
        assert [a.g.varix_to_edgelistix[v][k] for k in range(n)] == [b.g.varix_to_edgelistix[v][k] for k in range(n)]
    # per-factor parameter rows
    rad = pdm.prox_address("paradmm_prox_max_radius_3d")
    c = pdm.Graph(3)
    prm = np.array([[1.0], [2.0], [1.0]])
    c.add_factors(rad, np.array([[0], [1], [2]]), prm)
    pp = [C.cast(c.g.list_of_operators_param_pointers[f], C.POINTER(C.c_double))[0] for f in range(3)]
    assert pp == [1.0, 2.0, 1.0]
    with pytest.raises(pdm.ParadmmError):
        c.add_factors(rad, np.array([[-1]]))


def test_snapshot_round_trip(pdm, tmp_path):
    g = pdm.packing_square(10, 3, rho=7.0, alpha=0.5)
    rng = np.random.default_rng(1)
    for n in ("X", "M", "N", "U", "Z", "RHOS", "ALPHAS"):
        g.array(n)[:] = rng.uniform(-1, 1, size=g.array(n).shape)
    saved = {n: g.array(n).copy() for n in ("X", "M", "N", "U", "Z", "RHOS", "ALPHAS")}
    path = str(tmp_path / "state.pdm")
    g.save_state(path)
    for n in saved:
        g.array(n)[:] = 0.0
    g.load_state(path)
    for n in saved:
        assert np.array_equal(g.array(n), saved[n]), n
    # another structure, a truncated file and a foreign file are refused
    other = pdm.packing_square(11, 3)
    with pytest.raises(pdm.ParadmmError, match="another graph structure"):
        other.load_state(path)
    raw = open(path, "rb").read()
    # a truncated file (cut anywhere in the payload) or one with trailing
    # bytes is refused and leaves every array of the graph as it was
    marker = {n: rng.uniform(5, 6, size=g.array(n).shape) for n in saved}
    for n in saved:
        g.array(n)[:] = marker[n]
    for bad in (raw[:-100], raw[:len(raw) // 2], raw[:100], raw + b"\0" * 8):
        open(path, "wb").write(bad)
        with pytest.raises(pdm.ParadmmError, match="truncated"):
            g.load_state(path)
        for n in saved:
            assert np.array_equal(g.array(n), marker[n]), n
    open(path, "wb").write(b"not a snapshot at all" * 4)
    with pytest.raises(pdm.ParadmmError, match="not a snapshot"):
        g.load_state(path)


def _splitmix_unit(seed, count):
    """count uniforms in [0, 1) from splitmix64 (Steele, Lea, Flood 2014),
    53-bit mantissa: the stream paradmm_init_rand_seeded draws from."""
    x, out = seed & 0xFFFFFFFFFFFFFFFF, []
    for _ in range(count):
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        out.append((z >> 11) * (1.0 / 9007199254740992.0))
    return np.array(out)


def test_rand_init_seeded_draws_in_reference_order(pdm):
    """paradmm_init_rand_seeded: per entry X, N, M, U (each in its bounds),
    then Z, as initialize_X_N_Z_M_U_rand draws them (core/paradmm.c:247-261);
    the same seed gives the same state."""
    g = pdm.Graph(3)
    g.packing_square(5)
    b = dict(X=(-1.0, 2.0), N=(0.0, 0.5), Z=(3.0, 4.0), M=(-2.0, -1.0), U=(10.0, 12.0))
    g.init_rand_seeded(1234, **b)
    l1, l2 = 3 * g.nedge, 3 * g.nvar
    u = _splitmix_unit(1234, 4 * l1 + l2)
    for k, name in enumerate(("X", "N", "M", "U")):
        lo, hi = b[name]
        assert np.array_equal(g.array(name), lo + (hi - lo) * u[k:4 * l1:4]), name
    lo, hi = b["Z"]
    assert np.array_equal(g.array("Z"), lo + (hi - lo) * u[4 * l1:])
    z = g.array("Z").copy()
    g.init_rand_seeded(1234, **b)
    assert np.array_equal(g.array("Z"), z)
