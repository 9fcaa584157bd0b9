"""The reference demo packing/host.c, compiled unchanged against
include/paradmm.h and linked with libparadmm (oracle/_ref/hostc_link, built
where the reference sources exist), runs its own serial and OpenMP loops on
the GPU and matches the FP64 oracle on the same seeded initial state."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "oracle", "_ref", "hostc_link")


def test_hostc_link_built_unchanged():
    if not os.path.exists(EXE):
        pytest.skip("reference sources not present: oracle/_ref/hostc_link not built")
    out = subprocess.run(["nm", "-D", "--defined-only", EXE], capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    # host.c's operators are exported for the engine's by-name kind lookup
    for n in ("disk_collision_variable_radius", "disk_plane_collision_variable_radius",
              "maximize_sum_of_squared_disk_radii_alpha", "test_circlepacking_variableradius_triangle_auto"):
        assert n in names, n
    libs = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libparadmm.so" in libs


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["serial", "omp"])
def test_hostc_runs_on_gpu_and_matches_oracle(mode, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("reference sources not present: oracle/_ref/hostc_link not built")
    from oracle import pyoracle as po
    from gpu_helpers import trajectory_error
    nc, iters, seed = 8, 60, 3
    out = tmp_path / "z.bin"
    r = subprocess.run([EXE, mode, str(nc), str(iters), str(seed), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    z = np.fromfile(str(out)).reshape(iters, -1)
    o = po.Graph.triangle(nc, seed, -1.0, 2.0)
    hist = o.iterate(iters, record=True)
    err = trajectory_error(z, hist)
    print(f"host.c {mode}: Z error at it 1/10/30/60: " + " ".join(f"{err[k]:.1e}" for k in (0, 9, 29, 59)))
    assert err[:10].max() < 1e-5
