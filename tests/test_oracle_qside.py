"""The oracle's QSIDE weak heat-flux condition, pinned on the CPU.

QSIDE adds the integral of phi_i q over the chosen element sides to the
energy residual (src/bc_integ.c:1518-1519, 2156-2157, 2207), q being the heat
flux into the domain.  Pins:
  * on its own it adds q x (side length / face area) to the energy rows,
    spread over the side's nodes only, and nothing to other rows or to J;
  * conduction with T = 0 at x = 0 and QSIDE q at x = 1 (other walls
    adiabatic, no flow) solves to T = q x / k exactly at the nodes (the exact
    solution is linear, inside the isoparametric Q2 space), in 2D and 3D, on
    jittered meshes;
  * the flux into the domain is +q: reversing its sign reverses T;
  * bad sides and elements are refused.
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


@pytest.mark.parametrize("dim,n", [(2, 3), (3, 2)])
def test_qside_adds_q_times_side_measure(dim, n, oracle_mod):
    prob = _conduction(dim, n, 1.0, 0.15)
    nat = replace(prob, dbc_node=prob.dbc_node[:0], dbc_var=prob.dbc_var[:0],
                  dbc_value=prob.dbc_value[:0], datum_elem=-1)
    e, sd = exodus.wall_side_sets(nat.mesh)[2]  # x = 1 wall
    base = oracle_mod.Oracle(nat)
    x = np.random.default_rng(2).standard_normal(base.N) * 0.3  # QSIDE does not depend on x
    _, R0, V0 = base.fill(x)
    orc = oracle_mod.Oracle(nat)
    assert orc.add_qside(e, sd, 0.7) == 0
    _, R1, V1 = orc.fill(x)
    dR = R1 - R0
    assert np.array_equal(V1, V0)  # a constant flux has no Jacobian
    T = orc.unk_var == TEMP
    assert dR[T].sum() == pytest.approx(0.7 * 1.0, rel=1e-13)  # side length / face area 1
    assert not dR[~T].any()
    fu = orc.first_unknown.astype(np.int64)
    on_wall = np.zeros(nat.mesh.num_nodes, bool)
    on_wall[exodus.side_nodes(nat.mesh, e, sd).ravel()] = True
    tdof = fu[:-1] + dim
    assert not dR[tdof[~on_wall]].any() and np.all(dR[tdof[on_wall]] != 0)


@pytest.mark.parametrize("dim,n,k", [(2, 4, 2.5), (3, 2, 0.4)])
def test_flux_wall_conduction_is_linear_and_exact(dim, n, k, oracle_mod):
    prob = _conduction(dim, n, k, 0.15)
    e, sd = exodus.wall_side_sets(prob.mesh)[2]
    sols = []
    for q in (0.7, -0.7):
        orc = oracle_mod.Oracle(prob)
        assert orc.add_qside(e, sd, q) == 0
        x = _solve(oracle_mod, orc)
        tdof = orc.first_unknown.astype(np.int64)[:-1] + dim
        exact = q * prob.mesh.coor[0] / k
        assert np.abs(x[tdof] - exact).max() <= 1e-12 * abs(q / k)
        sols.append(x[tdof])
    assert np.array_equal(sols[0], -sols[1])


def test_qside_refuses_bad_input(oracle_mod):
    prob = _conduction(2, 2, 1.0, 0.0)
    orc = oracle_mod.Oracle(prob)
    assert orc.add_qside([0], [5], 1.0) == -1  # QUAD sides are 1..4
    assert orc.add_qside([prob.mesh.num_elems], [1], 1.0) == -1
    iso = oracle_mod.Oracle(problems.lid_cavity(n=2, Re=100.0))  # no energy equation
    assert iso.add_qside([0], [1], 1.0) == -1
    with pytest.raises(ValueError):
        orc.add_qside([0, 1], [1], 1.0)
