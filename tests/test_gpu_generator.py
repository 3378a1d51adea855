"""The device generator of the synthetic inputs (msod_box_mesh,
msod_flow_fields, msod_viscosity_fields) against the host generator it
mirrors (the Fortran host library): the same connectivity and random draws
exactly, coordinates and fields to the last bits of the libm functions, and
the same residuals when the kernels run on the device-generated arrays."""
import os
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_problem_on_host(n, p, **kw):
    """device_box_problem, read back and released."""
    d = m.device_box_problem(n, p, **kw)
    arrays = d.device_inputs() + (d.coord,)
    m.update_host(*arrays)
    m.delete(*arrays)
    return d


def compare(d, h):
    assert np.array_equal(d.connec, h.connec)
    assert np.abs(d.coord - h.coord).max() <= 1e-15
    assert rel_l2(d.He, h.He) <= 1e-13 and rel_l2(d.gpvol, h.gpvol) <= 1e-13
    for k in ("u", "q"):
        assert np.abs(getattr(d, k) - getattr(h, k)).max() <= 1e-14, k
    for k in ("rho", "pr", "E", "Tem"):
        assert rel_l2(getattr(d, k), getattr(h, k)) <= 1e-14, k
    assert np.array_equal(d.mu_fluid, h.mu_fluid)
    assert np.array_equal(d.mu_e, h.mu_e) and np.array_equal(d.mu_sgs, h.mu_sgs)


@pytest.mark.parametrize("p", range(1, 9))
def test_matches_host_generator_all_porders(gpu, p):
    say(f"[gpu] generator p={p}")
    kw = dict(ny=2, nz=3, jitter=0.05, rho_jump=0.1)
    compare(device_problem_on_host(3, p, **kw), m.box_problem(3, p, **kw))


def test_matches_host_generator_64cubed_p4(gpu):
    say("[gpu] generator 64^3 p=4 jittered")
    threads = min(16, os.cpu_count() or 1)
    compare(device_problem_on_host(64, 4, jitter=0.05),
            m.box_problem(64, 4, jitter=0.05, threads=threads))


def test_residuals_on_device_generated_inputs(gpu):
    """Kernels read the generated arrays in place (no upload)."""
    say("[gpu] residuals on device-generated inputs, 6x5x4 p=4")
    kw = dict(ny=5, nz=4, jitter=0.05)
    d = m.device_box_problem(6, 4, **kw)
    h = m.box_problem(6, 4, **kw)
    R, D = d.outputs(), d.outputs()
    m.create(*R, *D)
    try:
        m.full_convec_ijk(*d.convec_args(*R))
        m.full_diffusion_ijk(*d.diffu_args(*D))
        m.update_host(*R, *D)
    finally:
        m.delete(*R, *D, *d.device_inputs(), d.coord)
    Rh, Dh = h.outputs(), h.outputs()
    ins = h.device_inputs()
    m.copyin(*ins)
    m.create(*Rh, *Dh)
    try:
        m.full_convec_ijk(*h.convec_args(*Rh))
        m.full_diffusion_ijk(*h.diffu_args(*Dh))
        m.update_host(*Rh, *Dh)
    finally:
        m.delete(*ins, *Rh, *Dh)
    for a, b in zip(R + D, Rh + Dh):
        assert rel_l2(a, b) <= 1e-12


def test_arguments_are_checked(gpu):
    say("[gpu] generator argument checks")
    t = m.hex_tables(2)
    connec = np.empty((8, t.nnode), dtype=np.int32, order="F")
    coord = np.empty((125, 3), order="F")
    m.create(connec, coord)
    try:
        L = m.hip_lib()
        args = [m._ptr(a) for a in (t.xgp, t.AtoI, t.AtoJ, t.AtoK, connec, coord)]
        assert L.msod_box_mesh(2, 2, 2, 9, 0.0, 1, *args) == m.MSOD_ERR_PORDER
        assert L.msod_box_mesh(2, 2, 2, 2, 1.5, 1, *args) == m.MSOD_ERR_ARG
        assert L.msod_box_mesh(0, 2, 2, 2, 0.0, 1, *args) == m.MSOD_ERR_ARG
        # output arrays too small for a 3x2x2 box
        assert L.msod_box_mesh(3, 2, 2, 2, 0.0, 1, *args) == m.MSOD_ERR_MESH
        assert L.msod_box_mesh(2, 2, 2, 2, 0.0, 1, *args) == m.MSOD_OK
    finally:
        m.delete(connec, coord)
