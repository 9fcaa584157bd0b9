"""GPU tests of the contact state's spatial form against its id-order form,
through the C ABI.

The spatial form numbers the circles in the Morton order of their centres
when the state is entered and runs a contact tile only where the boxes of
its two blocks of circles meet (or its previous list holds pairs).  A
skipped tile holds no pair that can pass the prox's collision test, and
the circle sums are integers, so both forms give the same bits: Z and the
residuals inside the state, X and U after leaving it, and whole solves.
The id-order form is itself checked against the FP64 oracle
(test_gpu_contact.py, test_gpu_contact_full.py).
"""
import numpy as np
import pytest

from engine_amd import paradmm as P

pytestmark = pytest.mark.gpu

# iterate counts (runs of 64 and a remainder, single iterations) and leaves
PLAN = (1, 1, 64, 37, "leave", 70, "leave", 2)


def _run(monkeypatch, sparse, nc, seed, rho, plan=PLAN, tile="16"):
    monkeypatch.delenv("PARADMM_LAYOUT", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT_TILE", raising=False)
    monkeypatch.setenv("PARADMM_CONTACT_SPARSE", sparse)
    monkeypatch.setenv("PARADMM_SPARSE_TILE", tile)
    g = P.packing_square(nc, seed, rho, 1.0, 1.0)
    out = []
    for step in plan:
        if step == "leave":
            g.sync_to_host()
            out.append(("X", g.array("X").copy()))
            out.append(("U", g.array("U").copy()))
        else:
            g.iterate(step)
            out.append(("Z", g.read("Z")))
            out.append(("res", np.array(g.residuals())))
    info = g.contact_info()
    kern = set(g.profile_kernels(2))
    g.close()
    return out, info, kern


def _same(a, b, what):
    for (na, va), (nb, vb) in zip(a, b):
        assert na == nb
        assert np.array_equal(va, vb), f"{what}: {na} differs, max |diff| {np.max(np.abs(va - vb)):.3e}"


@pytest.mark.parametrize("nc,seed,rho,tile", [
    (20, 0, 1.0, "16"),      # one block pair
    (150, 1, 1.0, "16"),     # ragged last block
    (300, 0, 1.0, "32"),
    (2000, 2, 800.0, "16"),  # few contacts (rho = 0.4 N)
    (3000, 0, 1.0, "16"),
    (5000, 1, 1.0, "16"),
])
def test_spatial_form_equals_id_order(monkeypatch, nc, seed, rho, tile):
    ref, info_r, kern_r = _run(monkeypatch, "0", nc, seed, rho)
    got, info_g, kern_g = _run(monkeypatch, "1", nc, seed, rho, tile=tile)
    assert "contact_tile" in kern_r and "scs_rows" in kern_g and "contact_tile" not in kern_g
    assert info_g[1] == int(tile)
    # the same pairs are listed
    assert info_g[2] == info_r[2]
    _same(got, ref, f"N={nc} seed={seed} rho={rho}")
    print(f"N={nc}: spatial form equals the id-order form bit for bit, {info_g[2]} listed pairs")


def test_spatial_form_is_the_default_from_3200(monkeypatch):
    monkeypatch.delenv("PARADMM_CONTACT_SPARSE", raising=False)
    monkeypatch.delenv("PARADMM_CONTACT", raising=False)
    for nc, kern in ((3199, "contact_tile"), (3200, "scs_rows")):
        g = P.packing_square(nc, 0, 1.0, 1.0, 1.0)
        g.iterate(3)
        assert kern in set(g.profile_kernels(2))
        g.close()


def test_spatial_solve_equals_id_order_solve(monkeypatch):
    """A solve to tolerance (checks every 50 iterations, residuals and
    packing quality read inside the state) takes the same iterations to
    the same Z in both forms."""
    res = {}
    for sparse in ("0", "1"):
        monkeypatch.setenv("PARADMM_CONTACT_SPARSE", sparse)
        g = P.packing_square(3000, 3, 1200.0, 1.0, 1.0)
        r = g.solve(3000, check_every=50)
        res[sparse] = (r, g.read("Z"))
        g.close()
    r0, z0 = res["0"]
    r1, z1 = res["1"]
    assert r0["iterations"] == r1["iterations"]
    for k in ("converged", "primal_rms", "dual_rms", "max_overlap", "max_box", "objective"):
        assert r0[k] == r1[k], k
    assert np.array_equal(z0, z1)
    print(f"solve: {r1['iterations']} iterations, converged {r1['converged']}, primal {r1['primal_rms']:.3e}, "
          f"same bits in both forms")


@pytest.mark.parametrize("split,waves", [("1", "4"), ("3", "4"), ("8", "4"), ("3", "2"), ("2", "8"), ("1", "16")])
def test_spatial_launch_shapes_give_the_same_bits(monkeypatch, split, waves):
    """The row kernel's launch shape (workgroups per block row, wavefronts
    per workgroup) only shares the tiles out: every shape gives the bits of
    the default one (N = 3000, the default 5 parts of 4 wavefronts)."""
    monkeypatch.delenv("PARADMM_SPARSE_SPLIT", raising=False)
    monkeypatch.delenv("PARADMM_SPARSE_WAVES", raising=False)
    ref, info_r, _ = _run(monkeypatch, "1", 3000, 4, 1.0)
    monkeypatch.setenv("PARADMM_SPARSE_SPLIT", split)
    monkeypatch.setenv("PARADMM_SPARSE_WAVES", waves)
    got, info_g, kern = _run(monkeypatch, "1", 3000, 4, 1.0)
    assert "scs_rows" in kern
    assert info_g[2] == info_r[2]
    _same(got, ref, f"split {split}, waves {waves}")
