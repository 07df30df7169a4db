"""The HEX27 row kernel (one wave per node over its neighbours' column runs,
no column index per nonzero) against a sparse product of the same MSR matrix.

The kernel rebuilds every row's column list from the node graph: the
neighbours in ascending node order, their unknowns in order, the P columns
dropped from energy rows and the diagonal skipped.  A mistake in any of
these would misplace values, so the product is compared on lattice-ordered
and randomly renumbered meshes (where neighbour order and node order
differ), with and without the energy equation.  Bar: within 1e-13 of
max|A x| (the kernel and SciPy sum in different orders).  The residual form
(r = b - A x) runs inside every HEX27 multigrid test.
"""
from dataclasses import replace

import numpy as np
import pytest

from goma_amd import problems
from goma_amd.mesh import renumber

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def capi():
    from goma_amd import capi as M

    M.lib()
    assert torch.cuda.is_available()
    return M


def _renumbered(prob, seed):
    rng = np.random.default_rng(seed)
    m = prob.mesh
    node_order = rng.permutation(m.num_nodes)
    elem_order = rng.permutation(m.num_elems)
    new_id = np.empty_like(node_order)
    new_id[node_order] = np.arange(m.num_nodes)
    new_elem = np.empty_like(elem_order)
    new_elem[elem_order] = np.arange(m.num_elems)
    return replace(prob, mesh=renumber(m, node_order, elem_order),
                   dbc_node=new_id[prob.dbc_node].astype(np.int32),
                   datum_elem=int(new_elem[prob.datum_elem]))


CASES = {
    "energy": lambda: problems.lid_cavity(n=3, Re=400.0, dim=3, energy=True, Ra=1e4, jitter=0.1),
    "no_energy": lambda: problems.lid_cavity(n=3, Re=400.0, dim=3, jitter=0.1),
    "energy_renumbered": lambda: _renumbered(
        problems.lid_cavity(n=3, Re=400.0, dim=3, energy=True, Ra=1e4, jitter=0.1), 5),
    "no_energy_renumbered": lambda: _renumbered(problems.lid_cavity(n=3, Re=400.0, dim=3, jitter=0.1), 6),
}


@pytest.mark.parametrize("name", list(CASES))
def test_node_row_product_matches_sparse(name, capi, oracle_mod):
    prob = CASES[name]()
    print(f"[case] node row kernel {name}", flush=True)
    ctx, orc = capi.Context(prob), oracle_mod.Oracle(prob)
    rng = np.random.default_rng(13)
    x = rng.standard_normal(orc.N)
    rc, r, v = orc.fill(rng.standard_normal(orc.N) * 0.5)
    assert rc == 0
    A = oracle_mod.msr_to_scipy(orc.ija, v, orc.N).tocsr()
    want = A @ x
    vd = torch.from_numpy(v).cuda()
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(ctx.N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.matvec_dev(vd.data_ptr(), xd.data_ptr(), yd.data_ptr())
    got = yd.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    ctx.close()
