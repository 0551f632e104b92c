"""Per-routine parity of the HIP library (through the C ABI) against the C
oracle on seeded random chunks with ragged sizes.  Floating-point tolerance:
1e-11 relative to each field's scale (the kernels recompute Di and may
contract to FMA; the oracle never does)."""
import ctypes as C

import numpy as np
import pytest

import tl_oracle as O
from helpers import assert_close, assert_scalar, interior, random_chunk
from tealeaf_ref_amd import _lib as L

pytestmark = pytest.mark.gpu
TOL = 1e-11
STOL = 1e-10


class Pair:
    """The same chunk in the oracle and on the GPU."""

    def __init__(self, nx, ny, h, seed, coef=1, precond=1, reflective=False):
        self.nx, self.ny, self.h = nx, ny, h
        self.precond, self.reflective, self.coef = precond, reflective, coef
        self.rho, self.e1, self.vol, self.rx, self.ry = random_chunk(nx, ny, h, seed)
        self.o = O.OracleChunk(nx, ny, h)
        self.o.field("density")[:] = self.rho
        self.o.field("energy1")[:] = self.e1
        np.ctypeslib.as_array(self.o.c.volume, shape=self.vol.shape)[:] = self.vol
        self.o.set_rxry(self.rx, self.ry)
        self.d = L.Device(nx, ny, h)
        self.d.set_options(coef, precond, reflective)
        self.d.upload("density", self.rho)
        self.d.upload("energy1", self.e1)
        self.d.upload("volume", self.vol, halo=2)

    def close(self):
        self.d.close()

    def init(self):
        self.o.common_init(self.reflective, self.coef, self.precond)
        self.d.common_init(self.rx, self.ry)

    def cmp(self, name, tol=TOL, sl=None):
        a = self.d.download(name)
        b = self.o.field(name)
        if sl is None:
            a, b = interior(a, self.h), interior(b, self.h)
        else:
            a, b = a[sl], b[sl]
        assert_close(a, b, tol, name)


@pytest.fixture
def pair_factory():
    made = []

    def make(*a, **k):
        p = Pair(*a, **k)
        made.append(p)
        return p

    yield make
    for p in made:
        p.close()


@pytest.mark.parametrize("coef", [1, 2])
@pytest.mark.parametrize("precond", [1, 2, 3])
@pytest.mark.parametrize("reflective", [False, True])
def test_common_init(pair_factory, coef, precond, reflective):
    h = 1
    p = pair_factory(67, 45, h, seed=11 + coef + 3 * precond, coef=coef, precond=precond,
                     reflective=reflective)
    p.init()
    for f in ("u", "u0", "w", "r"):
        p.cmp(f)
    p.cmp("kx", sl=np.s_[1:, 1:])
    p.cmp("ky", sl=np.s_[1:, 1:])
    p.cmp("di", sl=np.s_[1:-1, 1:-1])
    if precond == 2:
        p.cmp("mi", sl=np.s_[1:-1, 1:-1])
    if precond == 3:
        for name, arr in (("cp", p.o.c.cp), ("bfp", p.o.c.bfp)):
            ref = np.ctypeslib.as_array(arr, shape=(p.ny, p.nx))
            assert_close(p.d.download(name, halo=0), ref, TOL, name)


@pytest.mark.parametrize("precond", [1, 2, 3])
def test_cg_routines(pair_factory, precond):
    p = pair_factory(130, 37, 1, seed=5, precond=precond)
    p.init()
    p.o.calc_residual()
    L.call("tl_calc_residual")
    p.cmp("r")
    assert_scalar(p.d.scalar("tl_calc_2norm", 1), p.o.calc_2norm(1), STOL, "2norm r")
    assert_scalar(p.d.scalar("tl_calc_2norm", 0), p.o.calc_2norm(0), STOL, "2norm u0")
    rro_o = p.o.cg_init(precond)
    rro_d = p.d.scalar("tl_cg_init")
    assert_scalar(rro_d, rro_o, STOL, "rro")
    p.cmp("p")
    p.cmp("z")
    rro = rro_o
    for it in range(4):
        pw_o = p.o.cg_calc_w()
        pw_d = p.d.scalar("tl_cg_calc_w")
        assert_scalar(pw_d, pw_o, STOL, f"pw it{it}")
        p.cmp("w")
        alpha = rro / pw_o
        rrn_o = p.o.cg_calc_ur(alpha, precond)
        rrn_d = p.d.scalar("tl_cg_calc_ur", alpha)
        assert_scalar(rrn_d, rrn_o, STOL, f"rrn it{it}")
        for f in ("u", "r"):
            p.cmp(f)
        beta = rrn_o / rro
        p.o.cg_calc_p(beta, precond, 0)
        L.call("tl_cg_calc_p", beta)
        p.cmp("p")
        rro = rrn_o


@pytest.mark.parametrize("precond", [1, 2, 3])
@pytest.mark.parametrize("reflective", [False, True])
def test_cheby_routines(pair_factory, precond, reflective):
    p = pair_factory(70, 66, 1, seed=7, precond=precond, reflective=reflective)
    p.init()
    fields_u = [0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0]
    p.o.update_halo(fields_u, 1, reflective)
    p.d.update_halo(fields_u, 1)
    p.o.calc_residual()
    L.call("tl_calc_residual")
    p.o.cg_init(precond)
    p.d.scalar("tl_cg_init")
    theta = 2.5
    p.o.cheby_init(theta, precond)
    L.call("tl_cheby_init", theta)
    for f in ("u", "p", "r", "w"):
        p.cmp(f)
    coefs = [(0.7, 0.3), (0.8, 0.25), (0.9, 0.2)]
    for a, b in coefs:
        p.o.cheby_iterate(a, b, precond)
        L.call("tl_cheby_iterate", a, b)
        for f in ("u", "p", "r", "w"):
            p.cmp(f)


@pytest.mark.parametrize("precond", [1, 2, 3])
@pytest.mark.parametrize("reflective", [False, True])
def test_ppcg_routines(pair_factory, precond, reflective):
    p = pair_factory(53, 71, 1, seed=9, precond=precond, reflective=reflective)
    p.init()
    p.o.calc_residual()
    L.call("tl_calc_residual")
    rro_o = p.o.ppcg_init(1, precond, 0)
    rro_d = p.d.scalar("tl_ppcg_init", 1)
    assert_scalar(rro_d, rro_o, STOL, "ppcg_init 1")
    p.cmp("p")
    L.call("tl_set_ppcg_active", 1)
    assert p.d.scalar("tl_ppcg_init", 2) == 0.0
    p.o.ppcg_init(2, precond, 1)
    p.cmp("z")
    theta = 3.0
    p.o.ppcg_init_sd(theta, precond)
    L.call("tl_ppcg_init_sd", theta)
    for f in ("sd", "rtemp", "utemp"):
        p.cmp(f)
    # the reference refreshes the (sd, r) halos before the first sweep
    p.o.update_halo([0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0], 1, reflective)
    al = np.array([0.5, 0.6, 0.7, 0.8])
    be = np.array([0.2, 0.3, 0.1, 0.4])
    sd_fields = [0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0]
    for i in range(4):
        p.o.ppcg_inner(al[i], be[i], precond)
        p.o.update_halo(sd_fields, 1, reflective)
    L.call("tl_ppcg_inner", al.ctypes.data_as(C.POINTER(C.c_double)),
           be.ctypes.data_as(C.POINTER(C.c_double)), 1, 4)
    for f in ("sd", "rtemp", "utemp"):
        p.cmp(f)
    p.o.ppcg_update_z()
    L.call("tl_ppcg_update_z")
    p.cmp("z")
    rro_o = p.o.ppcg_init(3, precond, 1)
    rro_d = p.d.scalar("tl_ppcg_init", 3)
    assert_scalar(rro_d, rro_o, STOL, "ppcg_init 3")
    p.cmp("p")
    # tea_solve.f90:250 refreshes p's halo before the next calc_w
    p.o.update_halo([0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0], 1, reflective)
    p.o.ppcg_store_r()
    L.call("tl_ppcg_store_r")
    p.cmp("r_store")
    zr_o = p.o.ppcg_calc_zrnorm(precond, 1)
    zr_d = p.d.scalar("tl_ppcg_calc_zrnorm")
    assert_scalar(zr_d, zr_o, STOL, "zrnorm")
    p.o.cg_calc_w()
    p.d.scalar("tl_cg_calc_w")
    p.o.cg_calc_ur(0.3, precond)
    p.d.scalar("tl_cg_calc_ur", 0.3)
    rrn_o = p.o.ppcg_calc_rrn()
    rrn_d = p.d.scalar("tl_ppcg_calc_rrn")
    assert_scalar(rrn_d, rrn_o, 1e-9, "rrn")


def test_jacobi_routine(pair_factory):
    for reflective in (False, True):
        p = pair_factory(61, 29, 1, seed=13, reflective=reflective)
        p.init()
        for it in range(3):
            e_o = p.o.jacobi_solve()
            e_d = p.d.scalar("tl_jacobi_solve")
            assert_scalar(e_d, e_o, STOL, "jacobi error")
            p.cmp("u")
            p.cmp("r")
        p.close()


@pytest.mark.parametrize("depth,h", [(1, 1), (1, 2), (2, 2)])
def test_update_halo(pair_factory, depth, h):
    p = pair_factory(33, 21, h, seed=3, reflective=True)
    rng = np.random.default_rng(1)
    names = ["density", "energy1", "u", "p", "sd", "rtemp", "kx"]
    for n in names:
        a = rng.standard_normal((p.ny + 2 * h, p.nx + 2 * h))
        p.o.field(n)[:] = a
        p.d.upload(n, a)
    fields = [1, 0, 1, 1, 1, 1, 1, 0, 1, 0, 0]  # FIELD_R slot refreshes rtemp
    p.o.update_halo(fields, depth, True)
    p.d.update_halo(fields, depth)
    for n in names:
        assert_close(p.d.download(n), p.o.field(n), 0.0, n)


def test_field_summary(pair_factory):
    p = pair_factory(300, 257, 2, seed=21)
    p.init()
    ref = p.o.field_summary()
    got = p.d.field_summary()
    for a, b, n in zip(got, ref, ("vol", "mass", "ie", "temp")):
        assert_scalar(a, b, 1e-12, n)
