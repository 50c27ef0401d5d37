"""CPU checks of the oracle's MPI driver (oracle/sia_mpi.c), the CPU baseline
of bench.py: PISM's 2-D block decomposition with periodic ghosts of width 2,
the gradient's ghost update, MPI_Allreduce(MAX) of D_max.

* Decomposition independence, as test/regression/test_02.sh checks it with a
  tolerance of 0: three steps of a window of the sheet give bitwise the same
  thickness on 1, 3 (3x1) and 4 (2x2) ranks.
* The single-rank run equals the oracle's serial steps (binding.step) on the
  same window generated by workloads.greenland_like, to 1e-12 of max H per
  step: the driver's C restatement of the generator agrees with numpy's to
  rounding.
"""
import os
import subprocess

import numpy as np
import pytest

import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "oracle", "sia_mpi")
MPIEXEC = "/opt/conda/bin/mpiexec"
WIN = dict(i0=150, j0=1370, wx=40, wy=48, Mz=101)

pytestmark = pytest.mark.skipif(not (os.path.exists(DRIVER) and os.path.exists(MPIEXEC)),
                                reason="MPICH or the MPI driver is not built")


def run(n, tmp_path, px=0, py=0, steps=3):
    out = os.path.join(str(tmp_path), "H_%d_%d_%d.bin" % (n, px, py))
    cmd = [MPIEXEC, "-n", str(n), DRIVER, "--i0", str(WIN["i0"]), "--j0", str(WIN["j0"]), "--wx", str(WIN["wx"]),
           "--wy", str(WIN["wy"]), "--Mz", str(WIN["Mz"]), "--steps", str(steps), "--warmup", "0",
           "--px", str(px), "--py", str(py), "--dump", out]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return np.fromfile(out, dtype=np.float64)


def test_decomposition_independent(tmp_path):
    H1 = run(1, tmp_path)
    H4 = run(4, tmp_path, 2, 2)
    H3 = run(3, tmp_path, 3, 1)
    assert H1.size == WIN["wx"] * WIN["wy"]
    assert np.array_equal(H1, H4) and np.array_equal(H1, H3)


def test_single_rank_equals_serial_oracle(oracle, tmp_path):
    w = W.greenland_like(WIN["i0"], WIN["i0"] + WIN["wx"], WIN["j0"], WIN["j0"] + WIN["wy"], Mz=WIN["Mz"])
    m = w.fields["mask"]
    assert (m == 0).any() and (m == 2).any()
    st = {k: w.fields[k].copy() for k in ("H", "b", "h", "mask")}
    Hmax = st["H"].max()
    for _ in range(3):
        oracle.step(w.grid, w.params, st, w.fields["E"], full=False)
    H1 = run(1, tmp_path)
    assert np.max(np.abs(H1 - st["H"])) <= 3e-12 * Hmax
