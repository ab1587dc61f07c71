"""The `mocc` command line on the device (src/driver.cpp:51-182,
eigen_solver.cpp:131-223): a deck with a <case_name> and <dump_iterations>,
run with command-line amendments. The program writes <case_name>.log,
<case_name>.h5 (with the amended deck as input_file) and one snapshot
<case_name>_iter_<n>.h5 per dump iteration; its eigenvalue is bitwise the
one the library's own solve gives for the amended options, and within 1e-5
of the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu

MOCC = os.path.join(ROOT, "mocc_amd", "lib", "mocc")


def test_cli_solve_output_and_snapshots(tmp_path):
    from mocc_amd import lattice_gen as lg
    from mocc_amd import output
    from mocc_amd.native import HostProblem
    from mocc_amd.sweeper import GpuMoCSweeper
    d = str(tmp_path)
    path = lg.write_config("pins3x3", d, name="c1_deck")
    text = open(path).read().replace('cmfd="f">', 'cmfd="f">\n  <dump_iterations>5 2</dump_iterations>')
    with open(path, "w") as f:
        f.write("<case_name>c1_cli</case_name>\n" + text)
    print("mocc CLI on the 3x3-pin deck", flush=True)
    r = subprocess.run([MOCC, "-a", "solver/k_tol=1e-7", "-a", "solver/psi_tol=1e-6", os.path.basename(path)],
                       cwd=d, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert "Convergence criteria satisfied!" in r.stdout
    assert "Output written to 'c1_cli'" in r.stdout
    for f in ("c1_cli.h5", "c1_cli.log", "c1_cli_iter_2.h5", "c1_cli_iter_5.h5"):
        assert os.path.exists(os.path.join(d, f)), f
    h5 = os.path.join(d, "c1_cli.h5")
    k_hist = output.read(h5, "convergence/k")
    deck = output.read_string(h5, "input_file")
    assert 'k_tol="1e-7"' in deck and 'psi_tol="1e-6"' in deck
    # the same solve through the library, with the amended tolerances
    hp = HostProblem.load(path)
    dev = GpuMoCSweeper(hp)
    rc, st = dev.eigen_solve(1e-7, 1e-6, hp.params().max_iter)
    assert rc == 0 and st.converged
    assert k_hist[-1] == st.k and len(k_hist) == st.iterations
    # snapshots hold the history up to their iteration
    for n in (2, 5):
        snap = output.read(os.path.join(d, f"c1_cli_iter_{n}.h5"), "convergence/k")
        assert np.array_equal(snap, k_hist[:n])
    ora = Oracle(hp, n_inner=1)
    rc_o, st_o, _ = ora.eigen_solve(1e-7, 1e-6, 1000)
    assert rc_o == 0 and abs(k_hist[-1] - st_o.k) < 1e-5
    assert "Running case: c1_cli" in open(os.path.join(d, "c1_cli.log")).read()
