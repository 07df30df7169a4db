"""Exodus II mesh / solution I/O (goma_amd.exodus over goma_amd.netcdf3), CPU only.

Pins:
  * the netCDF encoding against SciPy's independent netCDF-3 implementation
    (``scipy.io.netcdf_file``), both directions, classic and 64-bit-offset
    files, record and fixed variables, char arrays and typed attributes;
  * the Exodus names Goma's reader looks up (coordx.., connect1 with
    elem_type and 1-based nodes, eb_prop1 / ns_prop1 IDs, node_ns<k>,
    name_nod_var / vals_nod_var<i>, vals_elem_var<i>eb<b>) read through SciPy;
  * the dof map used to split a solution into Goma's fields against the
    oracle's own numbering;
  * a problem built from an Exodus file's node sets, with Dirichlet
    conditions given as deck cards, assembles to the same residual and
    Jacobian (oracle, bit for bit) as the structured constructor's problem;
  * a randomly renumbered mesh (nodes and elements) read back from Exodus
    solves to the structured problem's fields once the numbering is undone.
"""
import numpy as np
import pytest

from goma_amd import exodus, netcdf3 as nc, problems
from goma_amd.mesh import Mesh, renumber, structured_mesh
from goma_amd.problems import TEMP, U1, U2, U3

scipy_io = pytest.importorskip("scipy.io")


def _sample_vars():
    rng = np.random.default_rng(5)
    return {
        "x": rng.standard_normal(10),
        "c": np.arange(6, dtype=np.int32).reshape(2, 3),
        "sh": np.array([1, -2, 3], np.int16),
        "by": np.array([1, 2, 3, 4, 5], np.int8),
        "f": rng.standard_normal((2, 3)).astype(np.float32),
        "time": np.array([0.0, 0.5, 2.0]),
        "rec": rng.standard_normal((3, 5)),
        "names": nc.chars(["VX", "TEMP"], 7),
    }


@pytest.mark.parametrize("version", [1, 2])
def test_netcdf_written_here_reads_in_scipy(tmp_path, version):
    d = _sample_vars()
    p = tmp_path / "a.nc"
    dims = {"t": None, "n": 10, "a": 2, "b": 3, "m": 5, "k": 5, "len": 7}
    V = nc.Var
    nc.write(p, dims, [V("x", ("n",), d["x"], {"units": "m", "scale": 2.5}),
                       V("c", ("a", "b"), d["c"], {"elem_type": "QUAD9"}),
                       V("sh", ("b",), d["sh"]), V("by", ("k",), d["by"]), V("f", ("a", "b"), d["f"]),
                       V("time", ("t",), d["time"]), V("rec", ("t", "m"), d["rec"]),
                       V("names", ("a", "len"), d["names"])],
             {"api_version": np.float32(8.0), "title": "hello", "file_size": 1}, version=version)
    f = scipy_io.netcdf_file(p, "r", mmap=False)
    try:
        assert f.version_byte == version
        assert f.dimensions["t"] is None and f.dimensions["n"] == 10
        assert f._attributes["title"] == b"hello" and f._attributes["file_size"] == 1
        assert np.float32(f._attributes["api_version"]) == np.float32(8.0)
        for k in ("x", "c", "sh", "by", "f", "time", "rec"):
            assert np.array_equal(f.variables[k][:], d[k]), k
        assert f.variables["c"].elem_type == b"QUAD9"
        assert f.variables["x"].scale == 2.5
        assert [b"".join(r).decode() for r in f.variables["names"][:]] == ["VX", "TEMP"]
    finally:
        f.close()


@pytest.mark.parametrize("version", [1, 2])
def test_netcdf_written_by_scipy_reads_here(tmp_path, version):
    p = tmp_path / "b.nc"
    rng = np.random.default_rng(9)
    a, b = rng.standard_normal((3, 4)), rng.integers(-5, 5, 4).astype(np.int32)
    g = scipy_io.netcdf_file(p, "w", version=version)
    g.createDimension("time", None)
    g.createDimension("n", 4)
    g.title = b"written by scipy"
    va = g.createVariable("a", "d", ("time", "n"))
    va[:] = a
    vb = g.createVariable("b", "i", ("n",))
    vb[:] = b
    vb.name = b"ID"
    vz = g.createVariable("z", "b", ("time",))  # a byte record variable: record padding rules
    vz[:] = [1, 2, 3]
    vw = g.createVariable("w", "h", ("time",))
    vw[:] = [-7, 8, 9]
    g.close()
    ds = nc.read(p)
    assert ds.dims == {"time": 3, "n": 4} and ds.unlimited == "time"
    assert ds.attrs["title"] == "written by scipy"
    assert np.array_equal(ds.variables["a"].data, a)
    assert np.array_equal(ds.variables["b"].data, b) and ds.variables["b"].attrs["name"] == "ID"
    assert np.array_equal(ds.variables["z"].data, [1, 2, 3])
    assert np.array_equal(ds.variables["w"].data, [-7, 8, 9])


def test_netcdf_rejects_bad_input(tmp_path):
    p = tmp_path / "bad.nc"
    p.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(nc.NetCDFError, match="netCDF-4"):
        nc.read(p)
    good = tmp_path / "good.nc"
    nc.write(good, {"n": 1000}, [nc.Var("x", ("n",), np.arange(1000.0))])
    raw = good.read_bytes()
    (tmp_path / "cut.nc").write_bytes(raw[:len(raw) // 2])
    with pytest.raises(nc.NetCDFError, match="past the end"):
        nc.read(tmp_path / "cut.nc")
    (tmp_path / "cut2.nc").write_bytes(raw[:20])
    with pytest.raises(nc.NetCDFError, match="truncated"):
        nc.read(tmp_path / "cut2.nc")
    # corrupt headers: a negative dimension id, data offset or dimension length
    # must not wrap around in Python indexing
    import struct
    tiny = tmp_path / "tiny.nc"
    nc.write(tiny, {"n": 2}, [nc.Var("x", ("n",), np.arange(2.0))])
    hdr = bytearray(tiny.read_bytes())
    pos_dimid = hdr.index(b"x\0\0\0") + 8  # name "x" padded, then ndims, then the dim id
    bad = bytearray(hdr)
    bad[pos_dimid:pos_dimid + 4] = struct.pack(">i", -1)
    (tmp_path / "dimid.nc").write_bytes(bytes(bad))
    with pytest.raises(nc.NetCDFError, match="dimension id"):
        nc.read(tmp_path / "dimid.nc")
    bad = bytearray(hdr)
    pos_begin = pos_dimid + 4 + 8 + 8  # empty attribute list, type, vsize
    bad[pos_begin:pos_begin + 8] = struct.pack(">q", -8)
    (tmp_path / "begin.nc").write_bytes(bytes(bad))
    with pytest.raises(nc.NetCDFError, match="negative data offset"):
        nc.read(tmp_path / "begin.nc")
    bad = bytearray(hdr)
    pos_len = hdr.index(b"n\0\0\0") + 4
    bad[pos_len:pos_len + 4] = struct.pack(">i", -3)
    (tmp_path / "len.nc").write_bytes(bytes(bad))
    with pytest.raises(nc.NetCDFError, match="negative length"):
        nc.read(tmp_path / "len.nc")
    with pytest.raises(nc.NetCDFError, match="does not fit"):
        nc.write(p, {"n": 3}, [nc.Var("x", ("n",), np.zeros(4))])
    with pytest.raises(nc.NetCDFError, match="unknown dimension"):
        nc.write(p, {"n": 3}, [nc.Var("x", ("m",), np.zeros(3))])
    with pytest.raises(nc.NetCDFError, match="no netCDF classic type"):
        nc.write(p, {"n": 3}, [nc.Var("x", ("n",), np.zeros(3, np.uint32))])


@pytest.mark.parametrize("dim,n,jitter", [(2, 4, 0.15), (3, 2, 0.1)])
def test_exodus_mesh_round_trip_and_names(tmp_path, dim, n, jitter):
    mesh = structured_mesh(dim, n, jitter=jitter)
    sets = exodus.wall_node_sets(mesh)
    p = tmp_path / "mesh.exo"
    exodus.write(p, mesh, node_sets=sets, title="cavity")
    f = exodus.read(p)
    assert np.array_equal(f.mesh.coor, mesh.coor)  # FP64 bit for bit
    assert np.array_equal(f.mesh.connect, mesh.connect)
    assert sorted(f.node_sets) == sorted(sets)
    for k in sets:
        assert np.array_equal(f.node_sets[k], sets[k])
    assert f.title == "cavity" and f.block_ids == [1]
    # the names an Exodus reader (Goma's rd_exo.c through ex_get_*) looks up
    g = scipy_io.netcdf_file(p, "r", mmap=False)
    try:
        v = g.variables
        assert g.dimensions["num_dim"] == dim and g.dimensions["num_nodes"] == mesh.num_nodes
        assert g.dimensions["num_elem"] == mesh.num_elems and g.dimensions["num_el_blk"] == 1
        assert g.dimensions["num_nod_per_el1"] == (9 if dim == 2 else 27)
        assert v["connect1"].elem_type == (b"QUAD9" if dim == 2 else b"HEX27")
        assert v["connect1"][:].min() == 1 and v["connect1"][:].max() == mesh.num_nodes
        assert np.array_equal(v["coordx"][:], mesh.coor[0])
        assert v["eb_prop1"].name == b"ID" and v["ns_prop1"].name == b"ID"
        assert list(v["ns_prop1"][:]) == sorted(sets)
        assert np.array_equal(v["node_ns1"][:] - 1, sets[1])
        assert g._attributes["floating_point_word_size"] == 8
    finally:
        g.close()


def test_exodus_reads_older_layout_and_several_blocks(tmp_path):
    """``coord[num_dim][num_nodes]`` and ``vals_nod_var[time][var][node]`` (the
    pre-4.x layout), two element blocks, an element variable defined on one
    block only."""
    mesh = structured_mesh(2, 2)
    nn, ne = mesh.num_nodes, mesh.num_elems
    rng = np.random.default_rng(2)
    vals = rng.standard_normal((1, 2, nn))
    ev = rng.standard_normal((1, 3))
    V = nc.Var
    conn = mesh.connect.astype(np.int32) + 1
    dims = {"len_name": 33, "time_step": None, "num_dim": 2, "num_nodes": nn, "num_elem": ne,
            "num_el_blk": 2, "num_el_in_blk1": 3, "num_nod_per_el1": 9, "num_el_in_blk2": 1,
            "num_nod_per_el2": 9, "num_nod_var": 2, "num_elem_var": 1}
    nc.write(tmp_path / "old.exo", dims, [
        V("time_whole", ("time_step",), np.array([1.5])),
        V("eb_prop1", ("num_el_blk",), np.array([10, 20], np.int32), {"name": "ID"}),
        V("coord", ("num_dim", "num_nodes"), mesh.coor),
        V("connect1", ("num_el_in_blk1", "num_nod_per_el1"), conn[:3], {"elem_type": "QUAD"}),
        V("connect2", ("num_el_in_blk2", "num_nod_per_el2"), conn[3:], {"elem_type": "QUAD9"}),
        V("name_nod_var", ("num_nod_var", "len_name"), nc.chars(["VX", "VY"], 33)),
        V("vals_nod_var", ("time_step", "num_nod_var", "num_nodes"), vals),
        V("name_elem_var", ("num_elem_var", "len_name"), nc.chars(["P_E1"], 33)),
        V("elem_var_tab", ("num_el_blk", "num_elem_var"), np.array([[1], [0]], np.int32)),
        V("vals_elem_var1eb1", ("time_step", "num_el_in_blk1"), ev),
    ])
    f = exodus.read(tmp_path / "old.exo")
    assert np.array_equal(f.mesh.coor, mesh.coor) and np.array_equal(f.mesh.connect, mesh.connect)
    assert f.block_ids == [10, 20] and f.block_sizes == [3, 1]
    assert np.array_equal(f.times, [1.5])
    assert np.array_equal(f.nodal["VX"], vals[:, 0]) and np.array_equal(f.nodal["VY"], vals[:, 1])
    assert np.array_equal(f.elem["P_E1"][:, :3], ev) and np.isnan(f.elem["P_E1"][0, 3])


def test_exodus_rejects_other_element_types(tmp_path):
    mesh = structured_mesh(2, 1)
    bad = Mesh(dim=2, n=1, coor=mesh.coor, connect=mesh.connect[:, :4].copy())
    with pytest.raises(ValueError, match="QUAD9"):
        exodus.write(tmp_path / "q4.exo", bad)
    V = nc.Var
    nc.write(tmp_path / "q4.exo", {"num_dim": 2, "num_nodes": 9, "num_elem": 1, "num_el_blk": 1,
                                    "num_el_in_blk1": 1, "num_nod_per_el1": 4},
             [V("coordx", ("num_nodes",), mesh.coor[0]), V("coordy", ("num_nodes",), mesh.coor[1]),
              V("connect1", ("num_el_in_blk1", "num_nod_per_el1"), np.array([[1, 2, 3, 4]], np.int32),
                {"elem_type": "QUAD4"})])
    with pytest.raises(ValueError, match="QUAD9"):
        exodus.read(tmp_path / "q4.exo")


@pytest.mark.parametrize("name", ["cavity2d", "natconv2d", "cavity3d"])
def test_first_unknown_matches_oracle(name, oracle_mod):
    prob = {"cavity2d": lambda: problems.lid_cavity(n=3, Re=100.0),
            "natconv2d": lambda: problems.natural_convection(n=3, Ra=1e4, jitter=0.1),
            "cavity3d": lambda: problems.lid_cavity(n=2, dim=3, energy=True, Ra=1e3)}[name]()
    orc = oracle_mod.Oracle(prob)
    fu = exodus.first_unknown(prob.mesh, prob.phys.energy)
    assert np.array_equal(fu, orc.first_unknown.astype(np.int64))
    assert fu[-1] == orc.N


@pytest.mark.parametrize("dim", [2, 3])
def test_solution_written_in_goma_fields(tmp_path, dim, oracle_mod):
    prob = (problems.natural_convection(n=4, Ra=1e4) if dim == 2 else
            problems.lid_cavity(n=2, dim=3, energy=True, Ra=1e3))
    orc = oracle_mod.Oracle(prob)
    fu = orc.first_unknown.astype(np.int64)
    x = np.random.default_rng(4).standard_normal(orc.N)
    p = tmp_path / "soln.exo"
    exodus.write_solution(p, prob, x, time=0.0, node_sets=exodus.wall_node_sets(prob.mesh))
    f = exodus.read(p)
    names = ["VX", "VY", "VZ"][:dim] + ["T"]
    assert list(f.nodal) == names
    for a in range(dim):
        assert np.array_equal(f.nodal[names[a]][0], x[fu[:-1] + a])
    assert np.array_equal(f.nodal["T"][0], x[fu[:-1] + dim])
    cen = prob.mesh.connect[:, 8 if dim == 2 else 20]
    assert list(f.elem) == [f"P_E{k + 1}" for k in range(dim + 1)]
    for k in range(dim + 1):
        assert np.array_equal(f.elem[f"P_E{k + 1}"][0], x[fu[cen] + dim + 1 + k])
    # every unknown lands in exactly one output value
    total = sum(v.size for v in f.nodal.values()) + sum(v.size for v in f.elem.values())
    assert total == orc.N


def test_solution_size_is_checked():
    prob = problems.lid_cavity(n=2, Re=100.0)
    with pytest.raises(ValueError, match="unknowns"):
        exodus.solution_fields(prob, np.zeros(5))


def _deck_problem(tmp_path, prob, bcs):
    p = tmp_path / "deck_mesh.exo"
    exodus.write(p, prob.mesh, node_sets=exodus.wall_node_sets(prob.mesh))
    return exodus.problem_from_exodus(p, prob.phys, bcs, datum_elem=prob.datum_elem,
                                      datum_value=prob.datum_value)


def _assert_same_assembly(oracle_mod, a, b, seed=6):
    oa, ob = oracle_mod.Oracle(a), oracle_mod.Oracle(b)
    assert np.array_equal(oa.dbc_mask, ob.dbc_mask) and np.array_equal(oa.dbc_val, ob.dbc_val)
    x = np.random.default_rng(seed).standard_normal(oa.N) * 0.3
    ra, Ra, Va = oa.fill(x)
    rb, Rb, Vb = ob.fill(x)
    assert ra == rb == 0
    assert np.array_equal(Ra, Rb) and np.array_equal(Va, Vb)


def test_problem_from_exodus_node_sets_natural_convection(tmp_path, oracle_mod):
    """BC = U/V NS 1..4 0., BC = T NS 1 0.5, BC = T NS 2 -0.5 on the written mesh."""
    prob = problems.natural_convection(n=4, Ra=1e4, jitter=0.15)
    bcs = [(ns, var, 0.0) for var in (U1, U2) for ns in (1, 2, 3, 4)]
    bcs += [(1, TEMP, 0.5), (2, TEMP, -0.5)]
    _assert_same_assembly(oracle_mod, prob, _deck_problem(tmp_path, prob, bcs))


@pytest.mark.parametrize("dim", [2, 3])
def test_problem_from_exodus_lid_cavity_later_condition_holds(tmp_path, dim, oracle_mod):
    """The lid (ID 4) moves with u = 1; the side walls, listed after it, hold u = 0
    on the lid's rim (a later condition on the same unknown wins)."""
    prob = (problems.lid_cavity(n=4, Re=100.0) if dim == 2 else
            problems.lid_cavity(n=2, Re=100.0, dim=3, energy=True, Ra=1e3))
    vel = (U1, U2, U3)[:dim]
    walls = [k for k in range(1, 2 * dim + 1) if k != 4]
    bcs = [(4, U1, 1.0)] + [(4, v, 0.0) for v in vel[1:]]
    bcs += [(ns, v, 0.0) for ns in walls for v in vel]
    if dim == 3:
        bcs += [(4, TEMP, 0.5), (3, TEMP, -0.5)]
    _assert_same_assembly(oracle_mod, prob, _deck_problem(tmp_path, prob, bcs))


def test_problem_from_exodus_rejects_bad_cards(tmp_path):
    prob = problems.lid_cavity(n=2, Re=100.0)
    with pytest.raises(ValueError, match="node set 9"):
        _deck_problem(tmp_path, prob, [(9, U1, 0.0)])
    with pytest.raises(ValueError, match="datum"):
        _deck_problem(tmp_path, prob, [(1, 9, 0.0)])
    with pytest.raises(ValueError, match="energy"):
        _deck_problem(tmp_path, prob, [(1, TEMP, 0.0)])
    with pytest.raises(ValueError, match="not on this path"):
        _deck_problem(tmp_path, prob, [(1, U3, 0.0)])


def _renumbered_file(tmp_path, dim, seed=8):
    """The structured problem, and an Exodus file of its mesh with nodes and
    elements renumbered at random, node sets on the walls and the deck cards
    for its conditions.  Returns (prob, path, bcs, datum, pn, pe): the file's
    node ``pn[i]`` is node i, the file's element j is element ``pe[j]``."""
    prob = (problems.natural_convection(n=4, Ra=1e4, jitter=0.15) if dim == 2 else
            problems.lid_cavity(n=2, Re=100.0, dim=3, energy=True, Ra=1e3))
    rng = np.random.default_rng(seed)
    mesh = prob.mesh
    pn = rng.permutation(mesh.num_nodes)
    pe = rng.permutation(mesh.num_elems)
    coor = np.empty_like(mesh.coor)
    coor[:, pn] = mesh.coor
    ren = Mesh(dim=dim, n=0, coor=coor, connect=pn[mesh.connect[pe]].astype(np.int32))
    sets = {k: np.sort(pn[v]) for k, v in exodus.wall_node_sets(mesh).items()}
    p = tmp_path / "renumbered.exo"
    exodus.write(p, ren, node_sets=sets)
    vel = (U1, U2, U3)[:dim]
    if dim == 2:
        bcs = [(ns, v, 0.0) for v in vel for ns in (1, 2, 3, 4)] + [(1, TEMP, 0.5), (2, TEMP, -0.5)]
    else:
        walls = [k for k in range(1, 7) if k != 4]
        bcs = [(4, U1, 1.0)] + [(4, v, 0.0) for v in vel[1:]]
        bcs += [(ns, v, 0.0) for ns in walls for v in vel] + [(4, TEMP, 0.5), (3, TEMP, -0.5)]
    datum = int(np.flatnonzero(pe == prob.datum_elem)[0])
    return prob, p, bcs, datum, pn, pe


def _direct_newton(oracle_mod, prob):
    """Newton on the oracle's assembly with a SciPy direct solve (the linear
    solver is not what these tests are about; ILU(0) in a random order is a
    poor preconditioner).  Returns (x, oracle)."""
    spla = pytest.importorskip("scipy.sparse.linalg")
    orc = oracle_mod.Oracle(prob)
    x = np.zeros(orc.N)
    for it in range(20):
        rc, R, V = orc.fill(x)
        assert rc == 0
        if np.abs(R).max() < 1e-11:
            break
        x -= spla.spsolve(oracle_mod.msr_to_scipy(orc.ija, V, orc.N).tocsc(), R)
    assert np.abs(R).max() < 1e-11, f"no convergence: {np.abs(R).max()}"
    return x, orc


def _half_bandwidth(orc):
    rows = np.repeat(np.arange(orc.N), np.diff(orc.ija[:orc.N + 1]))
    cols = orc.ija[orc.ija[0]:orc.ija[orc.N]]
    return int(np.abs(cols - rows).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_renumbered_exodus_mesh_solves_to_the_same_fields(tmp_path, dim, oracle_mod):
    """Exodus files carry arbitrary node and element numbering.  Random
    renumbering of both, written and read back through Exodus, with the
    conditions given as node-set cards: the oracle's Newton solution, mapped
    back, equals the structured problem's field by field."""
    prob, p, bcs, datum, pn, pe = _renumbered_file(tmp_path, dim)
    rp = exodus.problem_from_exodus(p, prob.phys, bcs, datum_elem=datum, datum_value=prob.datum_value)
    assert rp.mesh.n == 0
    (na, ea), (nb, eb) = [exodus.solution_fields(q, _direct_newton(oracle_mod, q)[0]) for q in (prob, rp)]
    scale = max(np.abs(v).max() for v in na.values())
    for k in na:
        assert np.abs(nb[k][pn] - na[k]).max() <= 1e-9 * scale, k
    pscale = max(np.abs(v).max() for v in ea.values())
    for k in ea:
        assert np.abs(eb[k] - ea[k][pe]).max() <= 1e-9 * pscale, k


@pytest.mark.parametrize("dim", [2, 3])
def test_band_renumbering_narrows_the_band_and_writes_back_in_file_order(tmp_path, dim, oracle_mod):
    """``renumber=True``: the Jacobian's half bandwidth on the random file
    numbering is most of N; after ``band_order`` it is within 1.5x of the
    structured numbering's.  The solution written back lands in the file's
    numbering: read again, it equals the structured solution mapped by the
    file's permutation."""
    prob, p, bcs, datum, pn, pe = _renumbered_file(tmp_path, dim)
    raw = exodus.problem_from_exodus(p, prob.phys, bcs, datum_elem=datum, datum_value=prob.datum_value)
    rp = exodus.problem_from_exodus(p, prob.phys, bcs, datum_elem=datum, datum_value=prob.datum_value,
                                    renumber=True)
    x_ref, o_ref = _direct_newton(oracle_mod, prob)
    x, o = _direct_newton(oracle_mod, rp)
    bw_ref, bw_raw, bw = _half_bandwidth(o_ref), _half_bandwidth(oracle_mod.Oracle(raw)), _half_bandwidth(o)
    assert bw_raw > 0.5 * o.N
    assert bw <= 1.5 * bw_ref, (bw, bw_ref, bw_raw)
    out = tmp_path / "solution.exo"
    exodus.write_solution(out, rp, x)
    f = exodus.read(out)
    src = exodus.read(p)
    assert np.array_equal(f.mesh.coor, src.mesh.coor) and np.array_equal(f.mesh.connect, src.mesh.connect)
    assert all(np.array_equal(f.node_sets[k], src.node_sets[k]) for k in src.node_sets)
    na, ea = exodus.solution_fields(prob, x_ref)
    scale = max(np.abs(v).max() for v in na.values())
    for k in na:
        assert np.abs(f.nodal[k][0][pn] - na[k]).max() <= 1e-9 * scale, k
    pscale = max(np.abs(v).max() for v in ea.values())
    for k in ea:
        assert np.abs(f.elem[k][0] - ea[k][pe]).max() <= 1e-9 * pscale, k


def test_renumber_rejects_non_permutations():
    mesh = structured_mesh(2, 2)
    with pytest.raises(ValueError, match="node_order"):
        renumber(mesh, np.zeros(mesh.num_nodes, np.int64), np.arange(mesh.num_elems))
    with pytest.raises(ValueError, match="elem_order"):
        renumber(mesh, np.arange(mesh.num_nodes), np.zeros(mesh.num_elems, np.int64))


@pytest.mark.parametrize("dim,n", [(2, 8), (3, 3)])
def test_band_order_formula_and_box_recovery(dim, n, oracle_mod):
    """The per-element half-bandwidth formula equals the oracle's MSR
    pattern's (an energy problem: every unknown of two element-sharing nodes
    couples across the widest pair); on a randomly renumbered box without
    jitter the coordinate sweep gives back exactly the row-by-row band."""
    from goma_amd.mesh import _dof_half_bandwidth, band_order

    prob = (problems.natural_convection(n=n, Ra=1e4) if dim == 2 else
            problems.lid_cavity(n=n, dim=3, energy=True, Ra=1e3))
    mesh = prob.mesh
    ident = np.arange(mesh.num_nodes)
    assert _dof_half_bandwidth(mesh, ident) == _half_bandwidth(oracle_mod.Oracle(prob))
    pn = np.random.default_rng(12).permutation(mesh.num_nodes)
    coor = np.empty_like(mesh.coor)
    coor[:, pn] = mesh.coor
    rm = Mesh(dim=dim, n=0, coor=coor, connect=pn[mesh.connect].astype(np.int32))
    assert _dof_half_bandwidth(rm, np.arange(rm.num_nodes)) > 2 * _dof_half_bandwidth(mesh, ident)
    no, eo = band_order(rm)
    m2 = renumber(rm, no, eo)
    assert _dof_half_bandwidth(m2, np.arange(m2.num_nodes)) == _dof_half_bandwidth(mesh, ident)


# Node sets of each element side in Goma's get_side_info
# (src/mm_fill_fill.c:3049-3072 QUAD9, 3182-3250 HEX27), sides 1..
GOMA_SIDE_NODES = {
    2: [{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}],
    3: [{0, 8, 1, 13, 5, 16, 4, 12, 25}, {1, 9, 2, 14, 6, 17, 5, 13, 24},
        {2, 10, 3, 15, 7, 18, 6, 14, 26}, {0, 12, 4, 19, 7, 15, 3, 11, 23},
        {3, 10, 2, 9, 1, 8, 0, 11, 21}, {4, 16, 5, 17, 6, 18, 7, 19, 22}],
}


@pytest.mark.parametrize("dim", [2, 3])
def test_side_numbering_matches_goma(dim):
    npe = 9 if dim == 2 else 27
    one = Mesh(dim=dim, n=1, coor=np.zeros((dim, npe)), connect=np.arange(npe, dtype=np.int32)[None])
    for sid, want in enumerate(GOMA_SIDE_NODES[dim], start=1):
        assert set(exodus.side_nodes(one, [0], [sid])[0].tolist()) == want, sid
    with pytest.raises(ValueError, match="side numbers"):
        exodus.side_nodes(one, [0], [len(GOMA_SIDE_NODES[dim]) + 1])


@pytest.mark.parametrize("dim,n", [(2, 3), (3, 2)])
def test_side_sets_round_trip_and_cover_the_walls(tmp_path, dim, n):
    mesh = structured_mesh(dim, n, jitter=0.1)
    ss, ns = exodus.wall_side_sets(mesh), exodus.wall_node_sets(mesh)
    assert sorted(ss) == sorted(ns)
    for k, (e, sd) in ss.items():
        assert e.size == n ** (dim - 1)  # one side per boundary element on each wall
        nodes = exodus.side_nodes(mesh, e, sd)
        assert np.array_equal(np.unique(nodes), np.sort(ns[k]))
    p = tmp_path / "sides.exo"
    exodus.write(p, mesh, node_sets=ns, side_sets=ss)
    f = exodus.read(p)
    assert sorted(f.side_sets) == sorted(ss)
    for k in ss:
        assert np.array_equal(f.side_sets[k][0], ss[k][0]) and np.array_equal(f.side_sets[k][1], ss[k][1])
    g = scipy_io.netcdf_file(p, "r", mmap=False)
    try:
        assert g.dimensions["num_side_sets"] == len(ss) and g.variables["ss_prop1"].name == b"ID"
        assert np.array_equal(g.variables["elem_ss1"][:] - 1, ss[1][0])
        assert np.array_equal(g.variables["side_ss1"][:], ss[1][1])
    finally:
        g.close()
    with pytest.raises(ValueError, match="outside the mesh"):
        exodus.write(p, mesh, side_sets={1: (np.array([0]), np.array([7]))})


def test_multigrid_is_refused_on_a_mesh_without_hierarchy(tmp_path):
    """A mesh read from Exodus has no structured hierarchy: attaching the
    geometric multigrid raises before any library call."""
    from types import SimpleNamespace

    from goma_amd import capi

    prob = problems.natural_convection(n=2, Ra=1e4)
    p = tmp_path / "m.exo"
    exodus.write(p, prob.mesh, node_sets=exodus.wall_node_sets(prob.mesh))
    rp = exodus.problem_from_exodus(p, prob.phys, [(1, TEMP, 0.5)])
    with pytest.raises(ValueError, match="ILU"):
        capi.Context.attach_multigrid(SimpleNamespace(prob=rp))


def test_netcdf_append_record_reads_in_scipy(tmp_path):
    p = tmp_path / "rec.nc"
    rng = np.random.default_rng(4)
    a = rng.standard_normal((2, 5))
    nc.write(p, {"t": None, "m": 5}, [nc.Var("time", ("t",), np.array([0.0, 1.0])),
                                      nc.Var("a", ("t", "m"), a),
                                      nc.Var("fixed", ("m",), np.arange(5, dtype=np.int32))])
    b = rng.standard_normal(5)
    assert nc.append_record(p, {"time": 2.0, "a": b}) == 3
    with pytest.raises(nc.NetCDFError, match="exactly the record variables"):
        nc.append_record(p, {"a": b})
    with pytest.raises(nc.NetCDFError, match="shape"):
        nc.append_record(p, {"time": 3.0, "a": np.zeros(4)})
    f = scipy_io.netcdf_file(p, "r", mmap=False)
    try:
        assert np.array_equal(f.variables["time"][:], [0.0, 1.0, 2.0])
        assert np.array_equal(f.variables["a"][:], np.vstack([a, b]))
        assert np.array_equal(f.variables["fixed"][:], np.arange(5))
    finally:
        f.close()
    ds = nc.read(p)
    assert ds.dims["t"] == 3 and np.array_equal(ds.variables["a"].data, np.vstack([a, b]))


@pytest.mark.parametrize("renumber", [False, True])
def test_exodus_time_series_append(tmp_path, renumber, oracle_mod):
    """Goma writes each output step into one results file: steps appended to
    a solution file read back in order, in the file's numbering."""
    prob, p, bcs, datum, pn, pe = _renumbered_file(tmp_path, 2)
    rp = exodus.problem_from_exodus(p, prob.phys, bcs, datum_elem=datum, renumber=renumber)
    orc = oracle_mod.Oracle(rp)
    rng = np.random.default_rng(6)
    xs = [rng.standard_normal(orc.N) for _ in range(3)]
    out = tmp_path / "series.exo"
    exodus.write_solution(out, rp, xs[0], time=0.0)
    assert exodus.append_solution(out, rp, xs[1], 0.5) == 2
    assert exodus.append_solution(out, rp, xs[2], 1.0) == 3
    with pytest.raises(ValueError, match="does not follow"):
        exodus.append_solution(out, rp, xs[2], 1.0)
    f = exodus.read(out)
    assert np.array_equal(f.times, [0.0, 0.5, 1.0])
    for k, x in enumerate(xs):
        one = tmp_path / f"one{k}.exo"
        exodus.write_solution(one, rp, x, time=float(k))
        g = exodus.read(one)
        for name in g.nodal:
            assert np.array_equal(f.nodal[name][k], g.nodal[name][0])
        for name in g.elem:
            assert np.array_equal(f.elem[name][k], g.elem[name][0])
