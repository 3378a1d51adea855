"""CPU check of the assembly plan builder (minisod2d_amd/csrc/msod_plan.h,
the host code behind the kernels' atomic-free write phase): a small C++
checker compiled with AddressSanitizer and UBSan verifies every invariant the
kernels rely on, for box meshes of every order, shuffled element orders,
meshes with unreferenced nodes and the reference's disjoint mesh."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import minisod2d_amd as m

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "plan_check.cpp")
HDR = os.path.join(ROOT, "minisod2d_amd", "csrc", "msod_plan.h")
BIN = os.path.join(HERE, "_build", "plan_check")

# elements per batch of each order (Shape<P>::BATCH in csrc/msod_kernels.h)
BATCH = {1: 64, 2: 18, 3: 8, 4: 4, 5: 2, 6: 1, 7: 1, 8: 1}


@pytest.fixture(scope="module")
def checker():
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    newest = max(os.path.getmtime(SRC), os.path.getmtime(HDR))
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < newest:
        os.makedirs(os.path.dirname(BIN), exist_ok=True)
        subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=address,undefined",
                        "-fno-omit-frame-pointer", "-I", os.path.dirname(HDR), SRC, "-o", BIN],
                       check=True)
    return BIN


def run(checker, tmp_path, connec, npoin, l2g, B):
    nelem, nn = connec.shape
    path = str(tmp_path / "mesh.bin")
    with open(path, "wb") as fh:
        np.array([nelem, npoin], np.int64).tofile(fh)
        np.array([nn, B], np.int32).tofile(fh)
        np.asarray(l2g, np.int32).tofile(fh)
        np.asfortranarray(connec, dtype=np.int32).ravel(order="F").tofile(fh)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([checker, path], capture_output=True, text=True, env=env)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("p", range(1, 9))
def test_box_meshes(checker, tmp_path, p):
    n = {1: 9, 2: 6, 3: 5, 4: 4, 5: 3, 6: 3, 7: 2, 8: 2}[p]
    prob = m.box_problem(n, p, nz=n + 1)
    l2g = prob.tables.invAtoIJK.ravel(order="F") - 1
    run(checker, tmp_path, prob.connec, prob.npoin, l2g, BATCH[p])


@pytest.mark.parametrize("p", [1, 3, 4])
def test_shuffled_elements_and_unreferenced_nodes(checker, tmp_path, p):
    prob = m.box_problem(6 if p < 4 else 4, p)
    rng = np.random.default_rng(p)
    perm = rng.permutation(prob.nelem)
    keep = perm[: prob.nelem - 3]                     # three elements dropped
    connec = np.asfortranarray(prob.connec[keep])
    l2g = prob.tables.invAtoIJK.ravel(order="F") - 1
    out = run(checker, tmp_path, connec, prob.npoin, l2g, BATCH[p])
    assert "ncolors=" in out


@pytest.mark.parametrize("n,p", [(24, 1), (16, 3)])
def test_shuffled_meshes_past_32_colours_plan_atomic_only(checker, tmp_path, n, p):
    """Elements in random order need more than 32 batch colours: the plan is
    still built (the atomic assembly needs no colouring), without colours."""
    prob = m.box_problem(n, p)
    rng = np.random.default_rng(11)
    connec = np.asfortranarray(prob.connec[rng.permutation(prob.nelem)])
    l2g = prob.tables.invAtoIJK.ravel(order="F") - 1
    out = run(checker, tmp_path, connec, prob.npoin, l2g, BATCH[p])
    assert "atomic-only" in out, out


def test_disjoint_mesh_is_one_colour(checker, tmp_path):
    prob = m.reference_problem(10, 3)
    l2g = prob.tables.invAtoIJK.ravel(order="F") - 1
    out = run(checker, tmp_path, prob.connec, prob.npoin, l2g, BATCH[3])
    assert "ncolors=1 " in out


def test_out_of_range_connectivity_is_refused(checker, tmp_path):
    prob = m.box_problem(2, 2)
    connec = prob.connec.copy(order="F")
    connec[0, 0] = prob.npoin + 1
    nelem, nn = connec.shape
    path = str(tmp_path / "bad.bin")
    with open(path, "wb") as fh:
        np.array([nelem, prob.npoin], np.int64).tofile(fh)
        np.array([nn, BATCH[2]], np.int32).tofile(fh)
        (prob.tables.invAtoIJK.ravel(order="F") - 1).astype(np.int32).tofile(fh)
        connec.ravel(order="F").tofile(fh)
    r = subprocess.run([checker, path], capture_output=True, text=True)
    assert r.returncode == 2 and "outside 1..npoin" in r.stdout
