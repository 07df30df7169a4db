"""The oracle's QCONV convective heat-transfer condition, pinned on the CPU.

QCONV adds h (T_c - T) to the inward heat flux over the chosen element sides
(src/bc_integ.c:1492-1494: qrad_surf with zero emissivity,
src/mm_ns_bc.c:15145-15255), so the energy residual gains the side integral of
phi_i h (T_c - T) and the T-T Jacobian block -h times the side integral of
phi_i phi_j.  Pins:
  * conduction with T = 0 at x = 0 and QCONV (h, T_c) at x = 1 (other walls
    adiabatic, no flow) solves to T = h T_c x / (k + h) exactly at the nodes
    (a linear solution, inside the Q2 space), in 2D and 3D, on jittered
    meshes;
  * at a uniform temperature T0 it adds h (T_c - T0) x (side measure) to the
    energy rows of the side's nodes and nothing elsewhere;
  * its Jacobian equals the central difference of its residual (the term is
    linear in T, so the difference is exact up to rounding);
  * h = 0 changes nothing.
"""
from dataclasses import replace

import numpy as np
import pytest

from goma_amd import exodus, problems
from goma_amd.problems import TEMP

spla = pytest.importorskip("scipy.sparse.linalg")


def _conduction(dim, n, k, jitter):
    """No flow (g = 0), T = 0 on x = 0 only, walls no-slip."""
    prob = (problems.natural_convection(n=n, Ra=0.0, jitter=jitter) if dim == 2 else
            problems.lid_cavity(n=n, dim=3, energy=True, Ra=0.0, Re=1.0, jitter=jitter))
    keep = prob.dbc_var != TEMP
    lat = prob.mesh.lattice()
    x0 = np.flatnonzero(lat[:, 0] == 0)
    return replace(prob, dbc_node=np.concatenate([prob.dbc_node[keep], x0]).astype(np.int32),
                   dbc_var=np.concatenate([prob.dbc_var[keep], np.full(x0.size, TEMP)]).astype(np.int32),
                   dbc_value=np.concatenate([np.zeros(int(keep.sum())), np.zeros(x0.size)]),
                   phys=replace(prob.phys, kcond=k, g=(0.0, 0.0, 0.0)))


def _solve(oracle_mod, orc):
    x = np.zeros(orc.N)
    for _ in range(5):
        rc, R, V = orc.fill(x)
        assert rc == 0
        if np.abs(R).max() < 1e-12:
            break
        x -= spla.spsolve(oracle_mod.msr_to_scipy(orc.ija, V, orc.N).tocsc(), R)
    assert np.abs(R).max() < 1e-12
    return x


@pytest.mark.parametrize("dim,n,k,h,tc", [(2, 4, 2.5, 3.0, 0.8), (3, 2, 0.4, 1.5, -0.6)])
def test_convective_wall_conduction_is_linear_and_exact(dim, n, k, h, tc, oracle_mod):
    prob = _conduction(dim, n, k, 0.15)
    e, sd = exodus.wall_side_sets(prob.mesh)[2]  # x = 1 wall
    orc = oracle_mod.Oracle(prob)
    assert orc.add_qconv(e, sd, h, tc) == 0
    x = _solve(oracle_mod, orc)
    tdof = orc.first_unknown.astype(np.int64)[:-1] + dim
    exact = h * tc * prob.mesh.coor[0] / (k + h)
    assert np.abs(x[tdof] - exact).max() <= 1e-12 * abs(tc)


@pytest.mark.parametrize("dim,n", [(2, 3), (3, 2)])
def test_qconv_adds_h_dT_times_side_measure(dim, n, oracle_mod):
    prob = _conduction(dim, n, 1.0, 0.15)
    nat = replace(prob, dbc_node=prob.dbc_node[:0], dbc_var=prob.dbc_var[:0],
                  dbc_value=prob.dbc_value[:0], datum_elem=-1)
    e, sd = exodus.wall_side_sets(nat.mesh)[2]
    base = oracle_mod.Oracle(nat)
    x = np.random.default_rng(2).standard_normal(base.N) * 0.3
    T = base.unk_var == TEMP
    x[T] = 0.25  # uniform temperature
    _, R0, V0 = base.fill(x)
    orc = oracle_mod.Oracle(nat)
    assert orc.add_qconv(e, sd, 2.0, 1.25) == 0
    _, R1, V1 = orc.fill(x)
    dR = R1 - R0
    assert dR[T].sum() == pytest.approx(2.0 * (1.25 - 0.25) * 1.0, rel=1e-13)  # side measure 1
    assert not dR[~T].any()
    fu = orc.first_unknown.astype(np.int64)
    on_wall = np.zeros(nat.mesh.num_nodes, bool)
    on_wall[exodus.side_nodes(nat.mesh, e, sd).ravel()] = True
    tdof = fu[:-1] + dim
    assert not dR[tdof[~on_wall]].any() and np.all(dR[tdof[on_wall]] != 0)
    # h = 0: nothing
    zero = oracle_mod.Oracle(nat)
    assert zero.add_qconv(e, sd, 0.0, 1.25) == 0
    _, R2, V2 = zero.fill(x)
    assert np.array_equal(R2, R0) and np.array_equal(V2, V0)


@pytest.mark.parametrize("dim,n", [(2, 2), (3, 1)])
def test_qconv_jacobian_is_the_difference_of_its_residual(dim, n, oracle_mod):
    prob = _conduction(dim, n, 0.7, 0.1)
    e, sd = exodus.wall_side_sets(prob.mesh)[2]
    base, orc = oracle_mod.Oracle(prob), oracle_mod.Oracle(prob)
    h = 1.7
    assert orc.add_qconv(e, sd, h, 0.4) == 0
    x = np.random.default_rng(5).standard_normal(orc.N) * 0.3

    def term(v):
        _, Ra, Va = orc.fill(v)
        _, Rb, Vb = base.fill(v)
        return Ra - Rb, Va - Vb

    _, dV = term(x)
    J = oracle_mod.msr_to_scipy(orc.ija, dV, orc.N).toarray()
    tdof = np.flatnonzero(orc.unk_var == TEMP)
    eps = 1e-3
    for j in tdof:
        xp, xm = x.copy(), x.copy()
        xp[j] += eps
        xm[j] -= eps
        col = (term(xp)[0] - term(xm)[0]) / (2 * eps)
        assert np.abs(J[:, j] - col).max() <= 1e-9 * max(1.0, np.abs(col).max()), j
    assert np.abs(J).max() > 0.01 * h


def test_qconv_refuses_bad_input(oracle_mod):
    prob = _conduction(2, 2, 1.0, 0.0)
    orc = oracle_mod.Oracle(prob)
    assert orc.add_qconv([0], [5], 1.0, 0.0) == -1
    assert orc.add_qconv([prob.mesh.num_elems], [1], 1.0, 0.0) == -1
    iso = oracle_mod.Oracle(problems.lid_cavity(n=2, Re=100.0))
    assert iso.add_qconv([0], [1], 1.0, 0.0) == -1
